/*
 * Minimal re-declaration of qes::Domain (src/qes/Domain.h:41-196) for the
 * library's own tests of Solver_HIP: the members Solver and its subclasses
 * use, with the reference's semantics.  Not the reference's header; inside a
 * QES tree Solver_HIP.h compiles against the real one.
 */
#pragma once

#include <algorithm>
#include <vector>

namespace qes {

class Domain
{
public:
  /* Domain(nx, ny, nz, dx, dy, dz) takes cell counts and converts them to
   * the staggered sizes nx+1, ny+1, nz+2 with uniform dz (Domain.cpp:7-31). */
  Domain(int ncx, int ncy, int ncz, float dx, float dy, float dz)
    : nx_(ncx + 1), ny_(ncy + 1), nz_(ncz + 2), dx_(dx), dy_(dy), dz_(dz)
  {
    dz_array.assign(nz_ - 1, dz);
    defineVerticalGrid();
  }

  /* Stretched grid from the ncz layer thicknesses, dz_array[0] = dz_array[1]
   * (Domain::defineVerticalStretching(vector), Domain.cpp:133-144). */
  void setVerticalStretching(const std::vector<float> &dz_cells)
  {
    dz_array.assign(nz_ - 1, 0.f);
    for (int k = 1; k < nz_ - 1; ++k) dz_array[k] = dz_cells[k - 1];
    dz_array[0] = dz_array[1];
    dz_ = *std::min_element(dz_array.begin(), dz_array.end());
    defineVerticalGrid();
  }

  int nx() const { return nx_; }
  int ny() const { return ny_; }
  int nz() const { return nz_; }
  float dx() const { return dx_; }
  float dy() const { return dy_; }
  float dz() const { return dz_; }

  long numCellCentered() const { return (long)(nx_ - 1) * (ny_ - 1) * (nz_ - 1); }
  long numFaceCentered() const { return (long)nx_ * ny_ * nz_; }
  long cell(int i, int j, int k) const { return i + (long)j * (nx_ - 1) + (long)k * (nx_ - 1) * (ny_ - 1); }
  long face(int i, int j, int k) const { return i + (long)j * nx_ + (long)k * nx_ * ny_; }

  std::vector<float> dz_array;
  std::vector<float> z, z_face;

private:
  void defineVerticalGrid()
  {
    z_face.assign(nz_, 0.f);
    z_face[0] = -dz_array[0];
    z_face[1] = 0.0f;
    for (int k = 2; k < nz_; ++k) z_face[k] = z_face[k - 1] + dz_array[k - 1];
    z.assign(nz_ - 1, 0.f);
    z[0] = -0.5f * dz_array[0];
    for (int k = 1; k < nz_ - 1; ++k) z[k] = 0.5f * (z_face[k] + z_face[k + 1]);
  }

  int nx_, ny_, nz_;
  float dx_, dy_, dz_;
};

}  // namespace qes

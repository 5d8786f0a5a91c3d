/*
 * Minimal re-declaration of QESout::error (src/util/QESout.h:42-48): print
 * the message and exit, as the reference does.  Not the reference's header.
 */
#pragma once

#include <cstdio>
#include <cstdlib>
#include <string>

namespace QESout {
[[noreturn]] inline void error(std::string s)
{
  std::fprintf(stderr, "[ERROR]\t %s\n", s.c_str());
  std::exit(EXIT_FAILURE);
}
}  // namespace QESout

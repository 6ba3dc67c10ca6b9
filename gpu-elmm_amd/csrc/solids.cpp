// solids.cpp - solid-body masks and lists for axis-aligned cuboids.
#include "case_setup.hpp"
#include "errors.hpp"

namespace elmm {

int build_solids(const elmm_config &, const Grid1D &, int, const elmm_cuboid *, CaseData &) {
  return fail(ELMM_E_ARG, "solid bodies are not supported yet");
}

}  // namespace elmm

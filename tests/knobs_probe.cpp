// Prints the knobs multimergesort<T,cmp> runs with when this file is built
// with params.h macros on the command line (-DK=..., -DPIPELINE, ...), as the
// library resolves them: "K=<k> PL=<pl> pipeline=<kernel> blocks=<mergers>".
// Built and run by tests/test_knobs.py; host code only, no GPU needed.
#include <cstdio>

#include "gpu_mms/multimergesort.hxx"

int main() {
  const gmms_params p = gpu_mms::drop_in_params(0);
  gmms_params r;
  const int st = gmms_resolve_params(GMMS_INT32, &p, &r);
  if (st != GMMS_SUCCESS) {
    std::printf("resolve failed: %s\n", gmms_status_string(st));
    return 1;
  }
  std::printf("K=%d PL=%d pipeline=%d blocks=%d\n", r.kway, (int)PL, r.pipeline, p.mergers);
  return 0;
}

// Instantiates glynn31_kernel<NR, PHIP_GROUP> for every register-resident
// order NR in [kGlynnMinOrder, kGlynnMaxOrder].  The build compiles this file
// once per group size (0, 2, 3, 4, 6) so the instantiations spread over
// several translation units.
#include "launchers.h"
#include "glynn31.h"

#ifndef PHIP_GROUP
#error "PHIP_GROUP must be defined (0, 2, 3, 4 or 6)"
#endif

namespace phip {

namespace {
template <int NR>
hipError_t launch_nr(int nr, dim3 grid, hipStream_t stream, const Glynn31Args& a) {
    if constexpr (NR > kGlynnMaxOrder) {
        return hipErrorInvalidValue;
    } else {
        if (nr != NR) return launch_nr<NR + 1>(nr, grid, stream, a);
        hipLaunchKernelGGL((glynn31_kernel<NR, PHIP_GROUP>), grid, dim3(kGlynnBlock), 0, stream, a.tables,
                           a.primes, a.pinvs, a.n, a.cbits, a.lanes_per_prime, a.partials);
        return hipGetLastError();
    }
}
}  // namespace

#define PHIP_CAT2(a, b) a##b
#define PHIP_CAT(a, b) PHIP_CAT2(a, b)

hipError_t PHIP_CAT(launch_glynn31_g, PHIP_GROUP)(int nr, dim3 grid, hipStream_t stream,
                                                 const Glynn31Args& a) {
    if (nr < kGlynnMinOrder || nr > kGlynnMaxOrder) return hipErrorInvalidValue;
    return launch_nr<kGlynnMinOrder>(nr, grid, stream, a);
}

}  // namespace phip

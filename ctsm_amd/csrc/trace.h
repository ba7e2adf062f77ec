// Host-side trace ranges named after the reference's GPTL timers
// (perf_mod t_startf / t_stopf: clm_driver.F90:731 'canflux', :899
// 'soiltemperature', :920 'bgp2', :948 'hydro_without_drainage', :1146
// 'hydro2_drainage', :1413 'balchk'; CanopyFluxesMod.F90:1027 'can_iter';
// SoilTemperatureMod.F90:411 'SoilTempBandDiag', :654 'SoilThermProp',
// :1189 'PhaseChange').
//
// A range is pushed and popped through roctx around the launches of the
// routine it names.  The launches are asynchronous, so the range's own
// interval is the host's enqueue time; what it is for is the kernels: under
// `rocprofv3 --marker-trace --kernel-rename` every kernel dispatch is
// reported under the innermost range open when it was launched, so a kernel
// trace reads in the reference's timer names.  Without a profiler attached a
// push or pop is one call into libroctx that returns at once.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace ctsm {

class GptlRange {
public:
    explicit GptlRange(const char *name) { roctxRangePushA(name); }
    ~GptlRange() { roctxRangePop(); }
    GptlRange(const GptlRange &) = delete;
    GptlRange &operator=(const GptlRange &) = delete;
};

}  // namespace ctsm

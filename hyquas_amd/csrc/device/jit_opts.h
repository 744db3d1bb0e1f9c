// jit_opts.h -- hiprtc options of every specialised kernel.  Shared by the
// library (in-process compiles) and the compile helper hq_rtc, so both build
// the same code object from the same source.
#pragma once

namespace hq {

inline constexpr const char* kJitOpts[] = {"--offload-arch=gfx950", "-O3", "-ffp-contract=fast", "-Xclang",
                                           "-target-feature", "-Xclang", "-packed-fp32-ops"};
inline constexpr int kNumJitOpts = (int)(sizeof(kJitOpts) / sizeof(kJitOpts[0]));

}  // namespace hq

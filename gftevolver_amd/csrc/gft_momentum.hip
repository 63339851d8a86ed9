// gft_momentum.hip -- structure-of-arrays kick/drift form (placeholder until
// the fused kernel lands; every entry point reports "not available").
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "gft_math.hpp"

namespace gft {

struct MomentumState {
    void* base = nullptr;
};

inline int momentum_alloc(MomentumState&, const Coef&, bool) { return GFT_EINVAL; }
inline void momentum_free(MomentumState& m) {
    if (m.base) hipFree(m.base);
    m.base = nullptr;
}
template <class Src>
inline int momentum_import(MomentumState&, const Coef&, const Src*, const Src*, unsigned, hipStream_t, std::string& e) {
    e = "momentum form not available";
    return GFT_EINVAL;
}
template <class Dst>
inline int momentum_export(MomentumState&, const Coef&, Dst*, Dst*, unsigned, hipStream_t, std::string& e) {
    e = "momentum form not available";
    return GFT_EINVAL;
}
inline int momentum_step(MomentumState&, const Coef&, double, hipStream_t, unsigned*, std::string& e) {
    e = "momentum form not available";
    return GFT_EINVAL;
}
inline int momentum_gauss(MomentumState&, const Coef&, double*, long long, hipStream_t, std::string& e) {
    e = "momentum form not available";
    return GFT_EINVAL;
}

}  // namespace gft

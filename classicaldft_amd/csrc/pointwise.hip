// Launcher and instances of the specialised pointwise z pass (pw3.hpp), in a
// translation unit of their own so that they compile beside cdft_api.hip.
#include <type_traits>

#include "pw3.hpp"

namespace cdft {

namespace {

template <typename T, int MODEL, int M, int NA, bool MF>
long go(const PwArgs& a) {
  using G = Pw3Geo<T, M, NA + (MF ? 1 : 0)>;
  if constexpr (!G::fits) {
    return -1;
  } else {
    const long nblk = (a.nrows + G::RB - 1) / G::RB;
    hipLaunchKernelGGL((k_z_pw3<T, MODEL, M, NA, MF>), dim3((unsigned)nblk, a.members), dim3(G::NT), G::lds_bytes,
                       a.stream, static_cast<cplx<T>*>(a.B), a.arr_stride, a.vmf, a.nrows, a.Zp,
                       static_cast<const cplx<T>*>(a.twP), static_cast<const cplx<T>*>(a.twN), a.pp, a.partials,
                       a.flag, a.bt, a.ms);
    if (hipGetLastError() != hipSuccess) return -2;
    return nblk;
  }
}

template <typename T, int MODEL, int M>
long by_arrays(const PwArgs& a) {
  const bool mf = a.nconv > a.m;
  if constexpr (MODEL >= M_ESFMT) {
    if (a.m == 11 && !a.trace) return mf ? go<T, MODEL, M, 11, true>(a) : go<T, MODEL, M, 11, false>(a);
    if (a.m == 10 && a.trace) return mf ? go<T, MODEL, M, 10, true>(a) : go<T, MODEL, M, 10, false>(a);
  } else {
    if (a.m == 5) return mf ? go<T, MODEL, M, 5, true>(a) : go<T, MODEL, M, 5, false>(a);
  }
  return -1;
}

template <typename T, int M>
long by_model(const PwArgs& a) {
  switch (a.model) {
    case M_ROSENFELD: return by_arrays<T, M_ROSENFELD, M>(a);
    case M_RSLT: return by_arrays<T, M_RSLT, M>(a);
    case M_ESFMT: return by_arrays<T, M_ESFMT, M>(a);
    case M_WBI: return by_arrays<T, M_WBI, M>(a);
    case M_WBII: return by_arrays<T, M_WBII, M>(a);
  }
  return -1;
}

}  // namespace

// Launches k_z_pw3 for half-length M = Nz/2 when an instance covers (precision, M, model, arrays); returns the
// number of blocks launched, -1 when no instance covers the case (nothing launched), -2 on a launch error.
template <typename T>
long launch_pw3(const PwArgs& a, int M) {
  if (a.nconv < a.m || a.nconv > a.m + 1) return -1;
  switch (M) {
    case 128: return by_model<T, 128>(a);
    case 64: return by_model<T, 64>(a);
    case 48: return by_model<T, 48>(a);
    case 32: return by_model<T, 32>(a);
  }
  return -1;
}

template long launch_pw3<double>(const PwArgs&, int);
template long launch_pw3<float>(const PwArgs&, int);

}  // namespace cdft

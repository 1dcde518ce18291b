// m2l_tile.hip -- far field (M2L) with one lane per target node.
//
// Same operator as aran_development3d_m2l_rotate (arandevelopment3d.c:
// 423-430; rotate.c:165-207; translate.c:609-676), reorganised for a 64-wide
// wavefront: the targets of one level that sit in the same octant of their
// parent share the same 189 candidate source offsets, so a wave takes 64 such
// targets and walks the offsets in a fixed order; the waves of a workgroup
// take neighbouring tiles and walk the same offsets, so they share the
// per-offset tables in the scalar cache.  For each offset every
// lane translates its own source, while the rotation matrices (by theta),
// the phases (by phi) and the vertical factors (by |offset|) are the same
// for the whole wave and come through the scalar cache as SGPR operands.
// Each lane accumulates its target's local in registers, so the sum over
// sources has a fixed order and needs no atomics.  The 189 offsets of a tile
// are split over several waves (grid y); their partial locals are summed in
// a fixed order by the unscaling kernel.
//
// Charges are real, so every m = 0 coefficient of a multipole, of its
// rotations and of a local is real; their imaginary parts (zero up to
// rounding) are neither read nor computed, which also frees the registers
// that let two waves share a SIMD.
//
// Expansions are scaled per level by the box width w: M^_n = M_n / w^n,
// L^_l = L_l w^(l+1).  Then the vertical factor of offset o (box units) is
// (-1)^(l+m) c^(l+n) T(l,m,n) / |o|^(l+n+1), the same at every level.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "engine.h"
#include "m2l_tile.h"

namespace aran {

template <typename T> struct V2;
template <> struct V2<float> {
  typedef float2 type;
};
template <> struct V2<double> {
  typedef double2 type;
};

template <int P> struct Dims {
  static constexpr int MC = P * (P + 1) / 2;
  static constexpr int LC = (P + 1) * (P + 2) / 2;
  static constexpr int RB = (P + 1) * (P + 2) * (2 * P + 3) / 6;
  static constexpr int NV = (P + 1) * (P + 1) * P;  // [l][m][n]
};

constexpr __host__ __device__ int tri_c(int n) { return n * (n + 1) / 2; }

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}
constexpr __host__ __device__ int rofs_c(int l) {
  return l * (l + 1) * (2 * l + 1) / 6;
}

template <int P, typename T>
__global__ __launch_bounds__(64 * kTileWaves) __attribute__((
    amdgpu_waves_per_eu(m2l_tile_waves(P), m2l_tile_waves(P)))) void k_m2l_tile(
    const int *__restrict__ tile_src, const int *__restrict__ tile_oct,
    const short *__restrict__ oct_codes, const DirEntry *__restrict__ dir,
    const T *__restrict__ ephi, const T *__restrict__ rotT,
    const T *__restrict__ vfac, const typename V2<T>::type *__restrict__ Ms,
    typename V2<T>::type *__restrict__ Ls, int ntiles, int nsplit,
    int tile_major) {
  typedef typename V2<T>::type T2;
  constexpr int MC = Dims<P>::MC, LC = Dims<P>::LC, RB = Dims<P>::RB,
                VS = vfac_size(P);
  // Workgroups are dealt round-robin to the 8 XCDs; renumber them so that
  // each XCD (and its L2) walks one contiguous range of tile groups, with
  // all their offset ranges (tile_major) or one offset range at a time.
  const int ngroups = (ntiles + kTileWaves - 1) / kTileWaves;
  const int per_xcd = gridDim.x / 8;
  const int logical = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (logical >= ngroups * nsplit) return;
  const int split = tile_major ? logical % nsplit : logical / ngroups;
  const int group = tile_major ? logical / nsplit : logical % ngroups;
  // (the wave index is uniform per wave; readfirstlane tells the compiler)
  const int tile = group * kTileWaves +
                   __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int lane = threadIdx.x % 64;
  if (tile >= ntiles) return;  // whole wave; the kernel has no barrier
  // this wave takes offsets [k0, k1) of the tile; its partial local goes to
  // slice `split` of Ls (summed in slice order by the reduction kernel)
  const int chunk = (kOctOffsets + nsplit - 1) / nsplit;
  const int k0 = split * chunk;
  const int k1 = min(kOctOffsets, k0 + chunk);
  const int oct = __builtin_amdgcn_readfirstlane(tile_oct[tile]);
  T ar[LC], ai[LC];  // the lane's target local
  static_for<0, LC>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    ar[k] = 0;
    ai[k] = 0;
  });
  for (int k = k0; k < k1; k++) {
    const int B = tile_src[((size_t)tile * kOctOffsets + k) * 64 + lane];
    if (__ballot(B >= 0) == 0ull) continue;
    // wave-uniform operands: rotation matrices of theta, vertical factors
    // and phases of this offset, read through the scalar cache
    const int code = __builtin_amdgcn_readfirstlane(oct_codes[oct * kOctOffsets + k]);
    const int th = __builtin_amdgcn_readfirstlane(dir[code].theta);
    const T *__restrict__ sR = rotT + (size_t)th * 4 * RB;
    const T *__restrict__ sV = vfac + (size_t)code * VS;
    const T *__restrict__ sP = ephi + code * 2 * (P + 1);
    if (B >= 0) {
      T xr[MC], xi[MC];
      const T2 *mb = Ms + (size_t)B * MC;
      // phase: u_m = e^{i m phi} M_m
      static_for<0, P>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        static_for<0, l + 1>([&](auto mc) {
          constexpr int m = decltype(mc)::value;
          const T2 v = mb[tri_c(l) + m];
          if constexpr (m == 0) {
            xr[tri_c(l)] = v.x;
            xi[tri_c(l)] = 0;
          } else {
            const T c = sP[2 * m], s = sP[2 * m + 1];
            xr[tri_c(l) + m] = c * v.x - s * v.y;
            xi[tri_c(l) + m] = c * v.y + s * v.x;
          }
        });
      });
      // forward rotation, degree by degree: Re X = E+ Re u, Im X = E- Im u
      static_for<0, P>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        T tr[l + 1], ti[l + 1];
        static_for<0, l + 1>([&](auto mpc) {
          constexpr int mp = decltype(mpc)::value;
          const T *ep = sR + rofs_c(l) + mp * (l + 1);
          const T *em = ep + RB;
          T a = 0, b = 0;
          static_for<0, l + 1>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            a += ep[m] * xr[tri_c(l) + m];
            if constexpr (mp > 0 && m > 0) b += em[m] * xi[tri_c(l) + m];
          });
          tr[mp] = a;
          ti[mp] = b;
        });
        static_for<0, l + 1>([&](auto mpc) {
          constexpr int mp = decltype(mpc)::value;
          xr[tri_c(l) + mp] = tr[mp];
          xi[tri_c(l) + mp] = ti[mp];
        });
      });
      // vertical M2L and inverse rotation, degree by degree
      static_for<0, P + 1>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        T zr[l + 1], zi[l + 1];
        static_for<0, l + 1>([&](auto mc) {
          constexpr int m = decltype(mc)::value;
          T a = 0, b = 0;
          if constexpr (m < P) {
            const T *v = sV + vfac_ofs(P, l, m) - m;
            static_for<m, P>([&](auto nc) {
              constexpr int n = decltype(nc)::value;
              a += v[n] * xr[tri_c(n) + m];
              if constexpr (m > 0) b += v[n] * xi[tri_c(n) + m];
            });
          }
          zr[m] = a;
          zi[m] = b;
        });
        static_for<0, l + 1>([&](auto mpc) {
          constexpr int mp = decltype(mpc)::value;
          const T *fp = sR + 2 * RB + rofs_c(l) + mp * (l + 1);
          const T *fm = fp + RB;
          T wr = 0, wi = 0;
          static_for<0, l + 1>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            wr += fp[m] * zr[m];
            if constexpr (mp > 0 && m > 0) wi += fm[m] * zi[m];
          });
          if constexpr (mp == 0) {
            ar[tri_c(l)] += wr;
          } else {
            const T c = sP[2 * mp], s = sP[2 * mp + 1];
            ar[tri_c(l) + mp] += c * wr + s * wi;
            ai[tri_c(l) + mp] += c * wi - s * wr;
          }
        });
      });
    }
  }
  // slice layout [split][tile][coefficient][lane]: each store is one
  // contiguous 64-lane row
  T2 *out = Ls + ((size_t)split * ntiles + tile) * LC * 64 + lane;
  static_for<0, LC>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    T2 v;
    v.x = ar[k];
    v.y = ai[k];
    out[k * 64] = v;
  });
}

template <int P, typename T>
static void launch_one(int ntiles, const M2LTileArgs &a, const void *Ms,
                       void *Ls, hipStream_t s) {
  typedef typename V2<T>::type T2;
  const int ngroups = (ntiles + kTileWaves - 1) / kTileWaves;
  const int nblocks = (ngroups * a.nsplit + 7) / 8 * 8;
  k_m2l_tile<P, T><<<nblocks, 64 * kTileWaves, 0, s>>>(
      a.tile_src, a.tile_oct, a.oct_codes, a.dir,
      static_cast<const T *>(a.ephi), static_cast<const T *>(a.rot),
      static_cast<const T *>(a.vfac), static_cast<const T2 *>(Ms),
      static_cast<T2 *>(Ls), ntiles, a.nsplit, a.tile_major);
}

// Orders instantiated in this object (the build compiles this file once per
// range so that the instantiations compile in parallel).
#ifndef ARAN_M2L_PMIN
#define ARAN_M2L_PMIN 1
#define ARAN_M2L_PMAX 15
#endif

template <int P>
static hipError_t dispatch_range(int p, int ntiles, const M2LTileArgs &a,
                                 const void *Ms, void *Ls, hipStream_t s) {
  if constexpr (P > ARAN_M2L_PMAX) {
    return hipErrorInvalidValue;
  } else {
    if (p == P) {
      launch_one<P, float>(ntiles, a, Ms, Ls, s);
      return hipGetLastError();
    }
    return dispatch_range<P + 1>(p, ntiles, a, Ms, Ls, s);
  }
}

#define ARAN_CAT2(a, b) a##b
#define ARAN_CAT(a, b) ARAN_CAT2(a, b)
// launch_m2l_tile_<PMIN>: the FP32 tiled M2L for orders PMIN..PMAX
hipError_t ARAN_CAT(launch_m2l_tile_, ARAN_M2L_PMIN)(int p, int ntiles,
                                                     const M2LTileArgs &a,
                                                     const void *Ms, void *Ls,
                                                     hipStream_t s) {
  if (p < ARAN_M2L_PMIN || p > ARAN_M2L_PMAX) return hipErrorInvalidValue;
  return dispatch_range<ARAN_M2L_PMIN>(p, ntiles, a, Ms, Ls, s);
}

}  // namespace aran

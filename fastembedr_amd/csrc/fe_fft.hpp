// In-LDS Stockham FFT for one 256-thread workgroup (gfx950, wave64).
//
// A workgroup transforms NB independent length-N complex arrays that sit
// contiguously in LDS (array b at buf + b*N).  Radix-8 stages are used while
// at least 8 points remain per group, then one radix-4 or radix-2 stage.
// Each stage loads every butterfly input of the thread into registers,
// synchronises, applies the twiddles and the radix-R DFT, and writes the
// outputs back in place (Stockham autosort: the result is in natural order,
// no bit reversal).  Twiddles come from a table tw[t] = exp(-2*pi*i*t/N)
// computed on the host in double precision.
#pragma once

#include <hip/hip_runtime.h>

namespace fe {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  const float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  const float2 t2 = cadd(a1, a3), t3 = mul_mi<INV>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft8(float2 (&v)[8]) {
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  constexpr float r = 0.70710678118654752440f;
  // W8^1 = (1 -+ i)/sqrt2, W8^2 = -+i, W8^3 = (-1 -+ i)/sqrt2
  const float2 w1o = INV ? make_float2(r * (o1.x - o1.y), r * (o1.x + o1.y))
                         : make_float2(r * (o1.x + o1.y), r * (o1.y - o1.x));
  const float2 w2o = mul_mi<INV>(o2);
  const float2 w3o = INV ? make_float2(-r * (o3.x + o3.y), r * (o3.x - o3.y))
                         : make_float2(r * (o3.y - o3.x), -r * (o3.x + o3.y));
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, w1o);
  v[5] = csub(e1, w1o);
  v[2] = cadd(e2, w2o);
  v[6] = csub(e2, w2o);
  v[3] = cadd(e3, w3o);
  v[7] = csub(e3, w3o);
}

template <int R, bool INV>
__device__ __forceinline__ void dft_r(float2 (&v)[R]) {
  if constexpr (R == 2) {
    dft2<INV>(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4<INV>(v[0], v[1], v[2], v[3]);
  } else {
    dft8<INV>(v);
  }
}

template <int N, int NB, int NT, bool INV, int R, int NS>
__device__ __forceinline__ void fft_stage(float2* buf, const float2* __restrict__ tw) {
  constexpr int NBF = N / R;          // butterflies per array
  constexpr int TOT = NB * NBF;       // butterflies per stage
  constexpr int BPT = (TOT + NT - 1) / NT;
  constexpr int TSTEP = N / (NS * R);  // twiddle index stride
  float2 v[BPT][R];
  const int tid = static_cast<int>(threadIdx.x);
#pragma unroll
  for (int s = 0; s < BPT; ++s) {
    const int idx = tid + s * NT;
    if (TOT % NT == 0 || idx < TOT) {
      const int b = idx / NBF, j = idx - b * NBF;
      const float2* src = buf + b * N + j;
#pragma unroll
      for (int r = 0; r < R; ++r) v[s][r] = src[r * NBF];
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < BPT; ++s) {
    const int idx = tid + s * NT;
    if (TOT % NT == 0 || idx < TOT) {
      const int b = idx / NBF, j = idx - b * NBF;
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * k * TSTEP];
          if (INV) w.y = -w.y;
          v[s][r] = cmul(v[s][r], w);
        }
      }
      dft_r<R, INV>(v[s]);
      float2* dst = buf + b * N + (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * NS] = v[s][r];
    }
  }
  __syncthreads();
}

template <int N, int NB, int NT, bool INV, int NS>
__device__ __forceinline__ void fft_stages(float2* buf, const float2* __restrict__ tw) {
  if constexpr (NS < N) {
    constexpr int REM = N / NS;
    constexpr int R = REM >= 8 ? 8 : REM;
    fft_stage<N, NB, NT, INV, R, NS>(buf, tw);
    fft_stages<N, NB, NT, INV, NS * R>(buf, tw);
  }
}

// Unnormalised transform (forward: exp(-i), inverse: exp(+i)) of NB arrays of
// N points in LDS.  Caller must __syncthreads() after filling buf.
template <int N, int NB, int NT, bool INV>
__device__ __forceinline__ void fft_lds(float2* buf, const float2* __restrict__ tw) {
  fft_stages<N, NB, NT, INV, 1>(buf, tw);
}

}  // namespace fe

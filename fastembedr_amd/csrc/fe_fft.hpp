// In-LDS Stockham FFT for one workgroup of NT threads (gfx950, wave64).
//
// A workgroup transforms NB independent length-N complex arrays held in LDS
// with one padding slot after every 32 elements: element i of array b sits at
// buf[fft_at<N>(b, i)] (the padding spreads the strided accesses of the
// stages over the LDS banks).  Radix-8 stages are used while at least 8
// points remain per group, then one radix-4 or radix-2 stage.  Each stage
// loads every butterfly input of the thread into registers, synchronises,
// applies the twiddles and the radix-R DFT, and writes the outputs back in
// place (Stockham autosort: the result is in natural order, no bit reversal).
// Twiddles come from a table computed on the host in double precision, laid
// out stage by stage so that the R-1 twiddles of one butterfly are adjacent
// (16-byte loads); fft_tw_size(N) entries, filled by fft_twiddles().
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

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

constexpr int kFftPadShift = 5;  // one padding slot per 32 elements

// LDS float2 slots of one padded array of N elements
template <int N>
__host__ __device__ constexpr int fft_span() {
  return N + (N >> kFftPadShift);
}

// slot of element i of array b
template <int N>
__device__ __forceinline__ int fft_at(int b, int i) {
  return b * fft_span<N>() + i + (i >> kFftPadShift);
}

// radix of the stage whose butterflies span NS (radix 8 while 8 points remain)
__host__ __device__ constexpr int fft_radix(int n, int ns) { return n / ns >= 8 ? 8 : n / ns; }

// table slots per butterfly: R - 1 twiddles rounded up to an even count
__host__ __device__ constexpr int fft_tw_slots(int r) { return r == 8 ? 8 : (r == 4 ? 4 : 2); }

// first slot of the table of the stage with span NS (stages with NS > 1 only)
__host__ __device__ constexpr int fft_tw_offset(int n, int ns) {
  int off = 0;
  for (int s = 1; s < ns; s *= fft_radix(n, s))
    if (s > 1) off += s * fft_tw_slots(fft_radix(n, s));
  return off;
}

__host__ __device__ constexpr int fft_tw_size(int n) { return fft_tw_offset(n, n); }

// The table: for the stage with span NS and radix R, butterfly k in [0, NS)
// and r in [1, R): exp(-2*pi*i*r*k/(NS*R)) at fft_tw_offset(n, NS) + k *
// fft_tw_slots(R) + r - 1 (padding slots zero).  tw holds fft_tw_size(n).
inline void fft_twiddles(int n, float2* tw) {
  for (int s = 1; s < n; s *= fft_radix(n, s)) {
    if (s == 1) continue;
    const int r = fft_radix(n, s), sl = fft_tw_slots(r);
    float2* t = tw + fft_tw_offset(n, s);
    for (int k = 0; k < s; ++k) {
      for (int q = 0; q < sl; ++q) {
        const int rr = q + 1;
        if (rr < r) {
          const double a = -2.0 * 3.14159265358979323846 * static_cast<double>(rr) *
                           static_cast<double>(k) / (static_cast<double>(s) * r);
          t[k * sl + q] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
        } else {
          t[k * sl + q] = make_float2(0.f, 0.f);
        }
      }
    }
  }
}

template <int N, int NB, int NT, bool INV, int R, int NS>
__device__ __forceinline__ void fft_stage(float2* buf, const float2* __restrict__ tw) {
  constexpr int NBF = N / R;          // butterflies per array
  constexpr int TOT = NB * NBF;       // butterflies per stage
  constexpr int BPT = (TOT + NT - 1) / NT;
  constexpr int SL = fft_tw_slots(R);
  constexpr int NW = NS > 1 ? SL : 1;  // twiddle registers per butterfly
  float2 v[BPT][R];
  float2 w[BPT][NW];
  const int tid = static_cast<int>(threadIdx.x);
  // the butterfly inputs and, before the barrier, the twiddle loads
#pragma unroll
  for (int s = 0; s < BPT; ++s) {
    const int idx = tid + s * NT;
    if (TOT % NT == 0 || idx < TOT) {
      const int b = idx / NBF, j = idx - b * NBF;
#pragma unroll
      for (int r = 0; r < R; ++r) v[s][r] = buf[fft_at<N>(b, j + r * NBF)];
      if constexpr (NS > 1) {
        const float4* tp =
            reinterpret_cast<const float4*>(tw + fft_tw_offset(N, NS) + (j % NS) * SL);
#pragma unroll
        for (int q = 0; q < SL / 2; ++q) {
          const float4 t4 = tp[q];
          w[s][2 * q] = make_float2(t4.x, t4.y);
          w[s][2 * q + 1] = make_float2(t4.z, t4.w);
        }
      }
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
          float2 wr = w[s][r - 1];
          if (INV) wr.y = -wr.y;
          v[s][r] = cmul(v[s][r], wr);
        }
      }
      dft_r<R, INV>(v[s]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[fft_at<N>(b, base + r * NS)] = v[s][r];
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

// Unnormalised transform (forward: exp(-i), inverse: exp(+i)) of NB padded
// arrays of N points in LDS.  Caller must __syncthreads() after filling buf.
template <int N, int NB, int NT, bool INV>
__device__ __forceinline__ void fft_lds(float2* buf, const float2* __restrict__ tw) {
  fft_stages<N, NB, NT, INV, 1>(buf, tw);
}

}  // namespace fe

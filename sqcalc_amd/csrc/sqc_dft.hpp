// sqc_dft.hpp -- the small in-register DFTs of the line transforms (host and
// device).
//
// Every DFT here carries the +i sign of the reference transform
// (src/sqc_structure_factor.f90:1069-1071, iflag = +1):
// y_k = sum_r v_r exp(+2 pi i r k / R).
#pragma once
#include <hip/hip_runtime.h>

namespace sqc {

template <typename T> struct C2;
template <> struct C2<float> { using t = float2; };
template <> struct C2<double> { using t = double2; };

template <typename V> __host__ __device__ __forceinline__ V cadd(V a, V b) {
  return V{a.x + b.x, a.y + b.y};
}
template <typename V> __host__ __device__ __forceinline__ V csub(V a, V b) {
  return V{a.x - b.x, a.y - b.y};
}
template <typename V> __host__ __device__ __forceinline__ V cmul(V a, V b) {
  return V{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename V> __host__ __device__ __forceinline__ V cmul_i(V a) { return V{-a.y, a.x}; }
template <typename V, typename S> __host__ __device__ __forceinline__ V cscale(V a, S s) {
  return V{a.x * s, a.y * s};
}

template <typename T, int R> struct Dft;

template <typename T> struct Dft<T, 2> {
  using V = typename C2<T>::t;
  static __host__ __device__ __forceinline__ void run(V *v) {
    V a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T> struct Dft<T, 3> {
  using V = typename C2<T>::t;
  static __host__ __device__ __forceinline__ void run(V *v) {
    const T c = T(-0.5);
    const T s = T(0.86602540378443864676372317075294);
    V t1 = cadd(v[1], v[2]);
    V d = cscale(cmul_i(csub(v[1], v[2])), s);
    V m = cadd(v[0], cscale(t1, c));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, d);
    v[2] = csub(m, d);
  }
};

template <typename T> struct Dft<T, 4> {
  using V = typename C2<T>::t;
  static __host__ __device__ __forceinline__ void run(V *v) {
    V t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    V t2 = cadd(v[1], v[3]), t3 = cmul_i(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <typename T> struct Dft<T, 5> {
  using V = typename C2<T>::t;
  static __host__ __device__ __forceinline__ void run(V *v) {
    const T c1 = T(0.30901699437494742410229341718282);    // cos(2pi/5)
    const T c2 = T(-0.80901699437494742410229341718282);   // cos(4pi/5)
    const T s1 = T(0.95105651629515357211643933337938);    // sin(2pi/5)
    const T s2 = T(0.58778525229247312916870595463907);    // sin(4pi/5)
    V b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
    V d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    V t1 = cadd(v[0], cadd(cscale(b1, c1), cscale(b2, c2)));
    V t2 = cadd(v[0], cadd(cscale(b1, c2), cscale(b2, c1)));
    V u1 = cmul_i(cadd(cscale(d1, s1), cscale(d2, s2)));
    V u2 = cmul_i(csub(cscale(d1, s2), cscale(d2, s1)));
    v[0] = cadd(v[0], cadd(b1, b2));
    v[1] = cadd(t1, u1);
    v[4] = csub(t1, u1);
    v[2] = cadd(t2, u2);
    v[3] = csub(t2, u2);
  }
};

template <typename T> struct Dft<T, 8> {
  using V = typename C2<T>::t;
  static __host__ __device__ __forceinline__ void run(V *v) {
    const T r = T(0.70710678118654752440084436210485);
    V e[4] = {v[0], v[2], v[4], v[6]};
    V o[4] = {v[1], v[3], v[5], v[7]};
    Dft<T, 4>::run(e);
    Dft<T, 4>::run(o);
    // w8^k o_k, w8 = exp(i pi / 4)
    V w1 = V{(o[1].x - o[1].y) * r, (o[1].x + o[1].y) * r};
    V w2 = cmul_i(o[2]);
    V w3 = V{(-o[3].x - o[3].y) * r, (o[3].x - o[3].y) * r};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
  }
};

// DFT of 15 points by the prime-factor map 15 = 3 x 5: the input in the
// Ruritanian order n = (5 n1 + 3 n2) mod 15, the output in the Chinese
// remainder order k = (10 k1 + 6 k2) mod 15.  Then n k = 5 n1 k1 + 3 n2 k2
// (mod 15) and exp(2 pi i n k / 15) splits into the roots of 3 and of 5.
template <typename T> struct Dft<T, 15> {
  using V = typename C2<T>::t;
  static __host__ __device__ __forceinline__ void run(V *v) {
    V a[3][5];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 5; ++j) a[i][j] = v[(5 * i + 3 * j) % 15];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      V t[3] = {a[0][j], a[1][j], a[2][j]};
      Dft<T, 3>::run(t);
#pragma unroll
      for (int i = 0; i < 3; ++i) a[i][j] = t[i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) Dft<T, 5>::run(a[i]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 5; ++j) v[(10 * i + 6 * j) % 15] = a[i][j];
  }
};

}  // namespace sqc

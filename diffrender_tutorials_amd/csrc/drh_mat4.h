// drh_mat4.h -- LiteMath float4x4 arithmetic (common_src/LiteMath.h:1342-1634) for the host
// runtime and the kernels alike: one definition, so a matrix derived on the device (the
// transform update of the optimiser) is bit for bit the one the host derives at commit.
// Every translation unit is built with -ffp-contract=off and IEEE division.
#pragma once
#include <hip/hip_runtime.h>

namespace drh {

struct M4 {
  float m[16];  // column-major: m[c*4 + r]
  __host__ __device__ float operator()(int r, int c) const { return m[c * 4 + r]; }
  __host__ __device__ float& at(int r, int c) { return m[c * 4 + r]; }
};
__host__ __device__ inline M4 m4_load(const float* p) {
  M4 r;
  for (int i = 0; i < 16; i++) r.m[i] = p[i];
  return r;
}
__host__ __device__ inline M4 m4_identity() {
  M4 r{};
  r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
  return r;
}
__host__ __device__ inline void m4_vec(const M4& B, const float V[4], float R[4]) {
  for (int r = 0; r < 4; r++) R[r] = V[0] * B.m[r] + V[1] * B.m[4 + r] + V[2] * B.m[8 + r] + V[3] * B.m[12 + r];
}
__host__ __device__ inline M4 m4_mul(const M4& a, const M4& b) {
  M4 r;
  for (int c = 0; c < 4; c++) m4_vec(a, b.m + 4 * c, r.m + 4 * c);
  return r;
}
__host__ __device__ inline M4 m4_transpose(const M4& a) {
  M4 r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) r.m[j * 4 + i] = a.m[i * 4 + j];
  return r;
}

// Cofactor inverse with pair products, term for term as inverse4x4 (LiteMath.h:1551-1634).
__host__ __device__ inline M4 m4_inverse(const M4& a) {
  float p[12];
  M4 o;
  auto A = [&](int r, int c) { return a(r, c); };
  // cofactors of columns 0 and 1 from 2x2 products of rows 2, 3
  p[0] = A(2,2) * A(3,3); p[1] = A(2,3) * A(3,2); p[2] = A(2,1) * A(3,3); p[3] = A(2,3) * A(3,1);
  p[4] = A(2,1) * A(3,2); p[5] = A(2,2) * A(3,1); p[6] = A(2,0) * A(3,3); p[7] = A(2,3) * A(3,0);
  p[8] = A(2,0) * A(3,2); p[9] = A(2,2) * A(3,0); p[10] = A(2,0) * A(3,1); p[11] = A(2,1) * A(3,0);
  o.at(0,0) = p[0] * A(1,1) + p[3] * A(1,2) + p[4] * A(1,3);  o.at(0,0) -= p[1] * A(1,1) + p[2] * A(1,2) + p[5] * A(1,3);
  o.at(1,0) = p[1] * A(1,0) + p[6] * A(1,2) + p[9] * A(1,3);  o.at(1,0) -= p[0] * A(1,0) + p[7] * A(1,2) + p[8] * A(1,3);
  o.at(2,0) = p[2] * A(1,0) + p[7] * A(1,1) + p[10] * A(1,3); o.at(2,0) -= p[3] * A(1,0) + p[6] * A(1,1) + p[11] * A(1,3);
  o.at(3,0) = p[5] * A(1,0) + p[8] * A(1,1) + p[11] * A(1,2); o.at(3,0) -= p[4] * A(1,0) + p[9] * A(1,1) + p[10] * A(1,2);
  o.at(0,1) = p[1] * A(0,1) + p[2] * A(0,2) + p[5] * A(0,3);  o.at(0,1) -= p[0] * A(0,1) + p[3] * A(0,2) + p[4] * A(0,3);
  o.at(1,1) = p[0] * A(0,0) + p[7] * A(0,2) + p[8] * A(0,3);  o.at(1,1) -= p[1] * A(0,0) + p[6] * A(0,2) + p[9] * A(0,3);
  o.at(2,1) = p[3] * A(0,0) + p[6] * A(0,1) + p[11] * A(0,3); o.at(2,1) -= p[2] * A(0,0) + p[7] * A(0,1) + p[10] * A(0,3);
  o.at(3,1) = p[4] * A(0,0) + p[9] * A(0,1) + p[10] * A(0,2); o.at(3,1) -= p[5] * A(0,0) + p[8] * A(0,1) + p[11] * A(0,2);
  // columns 2 and 3 from 2x2 products of rows 0, 1
  p[0] = A(0,2) * A(1,3); p[1] = A(0,3) * A(1,2); p[2] = A(0,1) * A(1,3); p[3] = A(0,3) * A(1,1);
  p[4] = A(0,1) * A(1,2); p[5] = A(0,2) * A(1,1); p[6] = A(0,0) * A(1,3); p[7] = A(0,3) * A(1,0);
  p[8] = A(0,0) * A(1,2); p[9] = A(0,2) * A(1,0); p[10] = A(0,0) * A(1,1); p[11] = A(0,1) * A(1,0);
  o.at(0,2) = p[0] * A(3,1) + p[3] * A(3,2) + p[4] * A(3,3);  o.at(0,2) -= p[1] * A(3,1) + p[2] * A(3,2) + p[5] * A(3,3);
  o.at(1,2) = p[1] * A(3,0) + p[6] * A(3,2) + p[9] * A(3,3);  o.at(1,2) -= p[0] * A(3,0) + p[7] * A(3,2) + p[8] * A(3,3);
  o.at(2,2) = p[2] * A(3,0) + p[7] * A(3,1) + p[10] * A(3,3); o.at(2,2) -= p[3] * A(3,0) + p[6] * A(3,1) + p[11] * A(3,3);
  o.at(3,2) = p[5] * A(3,0) + p[8] * A(3,1) + p[11] * A(3,2); o.at(3,2) -= p[4] * A(3,0) + p[9] * A(3,1) + p[10] * A(3,2);
  o.at(0,3) = p[2] * A(2,2) + p[5] * A(2,3) + p[1] * A(2,1);  o.at(0,3) -= p[4] * A(2,3) + p[0] * A(2,1) + p[3] * A(2,2);
  o.at(1,3) = p[8] * A(2,3) + p[0] * A(2,0) + p[7] * A(2,2);  o.at(1,3) -= p[6] * A(2,2) + p[9] * A(2,3) + p[1] * A(2,0);
  o.at(2,3) = p[6] * A(2,1) + p[11] * A(2,3) + p[3] * A(2,0); o.at(2,3) -= p[10] * A(2,3) + p[2] * A(2,0) + p[7] * A(2,1);
  o.at(3,3) = p[10] * A(2,2) + p[4] * A(2,0) + p[9] * A(2,1); o.at(3,3) -= p[8] * A(2,1) + p[11] * A(2,2) + p[5] * A(2,0);
  const float k = 1.0f / (A(0,0) * o(0,0) + A(0,1) * o(1,0) + A(0,2) * o(2,0) + A(0,3) * o(3,0));
  for (int i = 0; i < 16; i++) o.m[i] = o.m[i] * k;
  return o;
}

// TransformR::to_mat (common_src/scene.h:99-104): translate4x4(t) * rotate4x4Z(rz) *
// rotate4x4Y(ry) * rotate4x4X(rx) * scale4x4(s), multiplied left to right, for
// x = (tx, ty, tz, rx, ry, rz, s) (LiteMath.h:1474-1519).
__host__ __device__ inline M4 m4_restricted(const float* x) {
  M4 T = m4_identity(), Rx = m4_identity(), Ry = m4_identity(), Rz = m4_identity(), S = m4_identity();
  T.m[12] = x[0]; T.m[13] = x[1]; T.m[14] = x[2];
  Rx.m[5] = cosf(x[3]); Rx.m[6] = sinf(x[3]); Rx.m[9] = -sinf(x[3]); Rx.m[10] = cosf(x[3]);
  Ry.m[0] = cosf(x[4]); Ry.m[2] = -sinf(x[4]); Ry.m[8] = sinf(x[4]); Ry.m[10] = cosf(x[4]);
  Rz.m[0] = cosf(x[5]); Rz.m[1] = sinf(x[5]); Rz.m[4] = -sinf(x[5]); Rz.m[5] = cosf(x[5]);
  S.m[0] = x[6]; S.m[5] = x[6]; S.m[10] = x[6];
  return m4_mul(m4_mul(m4_mul(m4_mul(T, Rz), Ry), Rx), S);
}

// Jacobian of the 3x4 row-major matrix (the order of the 12 transform gradient entries) with
// respect to x, closed form: the matrix is s * Rz Ry Rx in columns 0-2 and t in column 3
// (the function src/drender_autodiff.cpp:78-92 differentiates with Enzyme).  J[k][4r + c].
__host__ __device__ inline void restricted_jacobian(const float* x, float J[7][12]) {
  const float cx = cosf(x[3]), sx = sinf(x[3]), cy = cosf(x[4]), sy = sinf(x[4]);
  const float cz = cosf(x[5]), sz = sinf(x[5]), s = x[6];
  const float Rx[3][3] = {{1, 0, 0}, {0, cx, -sx}, {0, sx, cx}};
  const float Ry[3][3] = {{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}};
  const float Rz[3][3] = {{cz, -sz, 0}, {sz, cz, 0}, {0, 0, 1}};
  const float dRx[3][3] = {{0, 0, 0}, {0, -sx, -cx}, {0, cx, -sx}};
  const float dRy[3][3] = {{-sy, 0, cy}, {0, 0, 0}, {-cy, 0, -sy}};
  const float dRz[3][3] = {{-sz, -cz, 0}, {cz, -sz, 0}, {0, 0, 0}};
  auto mul3 = [](const float (*a)[3], const float (*b)[3], float (*o)[3]) {
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) o[r][c] = a[r][0] * b[0][c] + a[r][1] * b[1][c] + a[r][2] * b[2][c];
  };
  float YX[3][3], dYX[3][3], YdX[3][3], R[3][3], Drx[3][3], Dry[3][3], Drz[3][3];
  mul3(Ry, Rx, YX);
  mul3(Rz, YX, R);
  mul3(dRy, Rx, dYX);
  mul3(Ry, dRx, YdX);
  mul3(Rz, YdX, Drx);
  mul3(Rz, dYX, Dry);
  mul3(dRz, YX, Drz);
  for (int k = 0; k < 7; k++)
    for (int l = 0; l < 12; l++) J[k][l] = 0.0f;
  for (int r = 0; r < 3; r++) {
    J[r][4 * r + 3] = 1.0f;
    for (int c = 0; c < 3; c++) {
      J[3][4 * r + c] = s * Drx[r][c];
      J[4][4 * r + c] = s * Dry[r][c];
      J[5][4 * r + c] = s * Drz[r][c];
      J[6][4 * r + c] = R[r][c];
    }
  }
}

}  // namespace drh

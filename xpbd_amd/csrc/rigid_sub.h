// Buffers of one substep of the two-launch rigid substep (xpbd6_substep_*),
// shared by the kernels and the frame plan.
//
// desc.aux holds float4[6 n]: x_pred | q_pred | dx1 | drot1 | dv1 | dw1.
// Sums of substep s go to parity s & 1 of the Jacobi sums (parity 0:
// desc.dx / desc.drot, parity 1: dx1 / drot1) and of the velocity-pass sums
// (desc.dv / desc.dw, dv1 / dw1), its manifold points to pt_count[s & 1];
// each launch zeroes the parity that the next launch accumulates into, so no
// lane ever reads what another lane of the same launch writes.
#pragma once

#include <hip/hip_runtime.h>

#include "xpbd.h"

struct SubBufs {
  float4 *XPR, *QPR;          // predicted pose (written by the contacts launch)
  float4 *DX, *DR;            // this substep's Jacobi sums (zero on entry)
  float4 *DXn, *DRn;          // the next substep's (zeroed by the velocity launch)
  const float4 *DVin, *DWin;  // the previous substep's velocity-pass sums
  float4 *DV, *DW;            // this substep's (zeroed by the contacts launch)
  int *PTC, *PTCn;            // this / the next substep's point counter
  int apply;                  // apply DVin (every substep but a frame's first)
};

inline __host__ __device__ SubBufs sub_bufs(const xpbd_rigid_desc& d, int sub) {
  float4* aux = (float4*)d.aux;
  const size_t n = (size_t)d.n;
  float4* DX1 = aux + 2 * n;
  float4* DR1 = aux + 3 * n;
  float4* DV1 = aux + 4 * n;
  float4* DW1 = aux + 5 * n;
  const int a = sub & 1;
  SubBufs b;
  b.XPR = aux;
  b.QPR = aux + n;
  b.DX = a ? DX1 : (float4*)d.dx;
  b.DR = a ? DR1 : (float4*)d.drot;
  b.DXn = a ? (float4*)d.dx : DX1;
  b.DRn = a ? (float4*)d.drot : DR1;
  b.DV = a ? DV1 : (float4*)d.dv;
  b.DW = a ? DW1 : (float4*)d.dw;
  b.DVin = a ? (const float4*)d.dv : DV1;
  b.DWin = a ? (const float4*)d.dw : DW1;
  b.PTC = d.pt_count + a;
  b.PTCn = d.pt_count + (a ^ 1);
  b.apply = sub > 0;
  return b;
}

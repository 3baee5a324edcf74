// gp_serve.h — device interface of the serving loop's admission and
// retirement kernels.
#pragma once

#include "gp_internal.h"

namespace gp {

// Writable device arrays of one pool (a bucket whose lanes are slots).
struct PoolPtrs {
  int Bp, nb, nq, nv, nc, nh;
  int* kind;
  double* bp;
  float* bpf;
  double* cp;
  float* cpf;
  double* hp;
  float* hpf;
  double *q, *v, *tau, *vdot;
  int* fault;
  int* active;
};

// K admitted requests, packed request after request (device pointers).
struct AdmitBatch {
  int K;
  const int* lanes;   // [K]
  const int* kind;    // [K][nb]
  const int* steps;   // [K] horizon in steps
  const double* bp;   // [K][nb][NBF] (aligned-frame body parameters)
  const double* cp;   // [K][nc][NCF]
  const double* hp;   // [K][nh][NHF]
  const double* q;    // [K][nq]
  const double* v;    // [K][nv]
  const double* tau;  // [K][nv]
};

// K finished slots to pack (device pointers).
struct GatherBatch {
  int K;
  const int* lanes;  // [K]
  double* q;         // [K][nq]
  double* v;         // [K][nv]
  int* fault;        // [K]
};

int launch_admit(const PoolPtrs& p, const AdmitBatch& x, void* stream);
int launch_gather(const PoolPtrs& p, const GatherBatch& g, void* stream);

}  // namespace gp

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

// A registered model's parameters in device memory, uploaded once when the
// model is first submitted (device pointers).
struct ModelDev {
  const int* kind;   // [nb]
  const double* bp;  // [nb][NBF] (aligned-frame body parameters)
  const double* cp;  // [nc][NCF]
  const double* hp;  // [nh][NHF]
};

// K admitted requests, packed request after request (device pointers).
struct AdmitBatch {
  int K;
  const int* lanes;   // [K]
  const int* model;   // [K] index into the table's models
  const int* steps;   // [K] horizon in steps
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

// Every pool's admissions (or finished slots) of one serving round, for one
// launch: table entry e covers requests [start[e], start[e + 1]).
struct AdmitTable {
  int n;
  const ModelDev* models;  // the server's registered models
  int start[kMaxGroup + 1];
  PoolPtrs p[kMaxGroup];
  AdmitBatch x[kMaxGroup];
};
struct GatherTable {
  int n;
  int start[kMaxGroup + 1];
  PoolPtrs p[kMaxGroup];
  GatherBatch g[kMaxGroup];
};

// One workgroup per request; its lanes copy the request's values in parallel.
// d_table and everything it points to are device memory.
int launch_admit(const AdmitTable* d_table, int n_requests, void* stream);
int launch_gather(const GatherTable* d_table, int n_requests, void* stream);

}  // namespace gp

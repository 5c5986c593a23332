// host_step.hip -- the device step (cable_amd/csrc/cbm_kernel.h, step_tile)
// compiled for the CPU, so that the test suite can run the GPU's physics on
// host arrays under the sanitizers and compare it with the oracle
// (tests/test_host_step.py builds and drives it).
//
//   host_step IN OUT
//
// IN (little-endian): int64 mp, int32 nsteps, int32 full (1: every field in
// and out, the drop-in step; 0: the resident run's compact set with the
// class tables; 2: the resident set stepped by the hybrid method, mixed
// precision only, from a fixed warm start -- leaf temperature tk + 0.5 K,
// leaf-surface CO2 0.7 ca -- with one polish pass and no residual allowance,
// so that many tiles fall back and start again), int32 precision (0 FP64,
// 1 mixed), double dels, 18 doubles
// zse(6) zshh(7) ratecp(3) ratecs(2), the veg and soil class tables
// (17 x cable_gpu_veg_cols, 9 x cable_gpu_soil_cols doubles, in the column
// order of include/cable_gpu.h), then per step: int32 ktau, int32
// first_call and every field of include/cable_fields.h in table order as
// (n, mp) doubles -- the state the step starts from.  OUT: per step, every
// field after the step, as doubles in the same layout.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../cable_amd/csrc/cbm_kernel.h"

using namespace cbm;

namespace {
const int NEL[] = {
#define X(id, name, prec, n, role) n,
    CABLE_FIELDS(X)
#undef X
};
const char PREC[] = {
#define X(id, name, prec, n, role) #prec[0],
    CABLE_FIELDS(X)
#undef X
};

bool rd(FILE *f, void *p, size_t n) { return fread(p, 1, n, f) == n; }

// device element type of field i for working precision R: 0 double, 1 float, 2 int
template <class R>
int dev_kind(int i) {
  if (PREC[i] == 'I') return 2;
  if (PREC[i] == 'D' || PREC[i] == 'B') return 0;
  return sizeof(R) == 8 ? 0 : 1;
}

template <class R>
int run(FILE *in, FILE *out, int64_t mp, int nsteps, int full, double dels, const double *glob,
        const std::vector<double> &vtab, const std::vector<double> &stab) {
  if (full == 2 && sizeof(R) != 4) return 4;
  long long fallbacks = 0;
  std::vector<std::vector<double>> h(CABLE_NFIELDS);
  std::vector<void *> dev(CABLE_NFIELDS);
  for (int i = 0; i < CABLE_NFIELDS; i++) {
    h[i].resize((size_t)NEL[i] * mp);
    const int k = dev_kind<R>(i);
    dev[i] = malloc((size_t)NEL[i] * mp * (k == 0 ? 8 : 4));
  }
  StepArgs a;
  memset(&a, 0, sizeof a);
  a.F = dev.data();
  a.mp = mp;
  a.driver = 1;
  a.dels = dels;
  memcpy(a.zse, glob, sizeof a.zse);
  memcpy(a.zshh, glob + 6, sizeof a.zshh);
  memcpy(a.ratecp, glob + 13, sizeof a.ratecp);
  memcpy(a.ratecs, glob + 16, sizeof a.ratecs);
  a.vtab = vtab.data();
  a.stab = stab.data();
  // the hybrid method's inputs (full == 2)
  std::vector<float> pred((size_t)8 * mp, 0.0f), hyb((size_t)3 * mp, 0.0f);
  for (int64_t t = 0; t < mp; t++) {
    pred[8 * t + 0] = 0.5f;
    pred[8 * t + 1] = pred[8 * t + 2] = 0.7f;
  }
  a.pred = pred.data();
  a.hyb = hyb.data();
  a.max_passes = 1;
  a.resid_max = 0.0f;
  for (int s = 0; s < nsteps; s++) {
    int32_t ktau = 0, first = 0;
    if (!rd(in, &ktau, 4) || !rd(in, &first, 4)) return 2;
    for (int i = 0; i < CABLE_NFIELDS; i++) {
      if (!rd(in, h[i].data(), sizeof(double) * h[i].size())) return 2;
      const int k = dev_kind<R>(i);
      for (size_t j = 0; j < h[i].size(); j++) {
        if (k == 0) ((double *)dev[i])[j] = h[i][j];
        else if (k == 1) ((float *)dev[i])[j] = (float)h[i][j];
        else ((int *)dev[i])[j] = (int)h[i][j];
      }
    }
    a.ktau = ktau;
    a.first_call = first;
    // the resident run's scalar forcing (uniform over tiles)
    a.ca = h[CABLE_F_met_ca][0];
    a.doy = h[CABLE_F_met_doy][0];
    for (int64_t t = 0; t < mp; t++) {
      if (full == 1) {
        TileOut<MODE_PLAIN, false> o;
        step_tile<R, true, false, MODE_PLAIN, false>(a, t, nullptr, nullptr, o);
      } else if (full == 0) {
        TileOut<MODE_PLAIN, false> o;
        step_tile<R, false, true, MODE_PLAIN, false>(a, t, vtab.data(), stab.data(), o);
      } else if constexpr (sizeof(R) == 4) {
        TileOut<MODE_HYBRID, false> o;
        step_tile<R, false, true, MODE_HYBRID, false>(a, t, vtab.data(), stab.data(), o);
        fallbacks += o.fallback ? 1 : 0;
      }
    }
    a.hyb_valid = 1;
    for (int i = 0; i < CABLE_NFIELDS; i++) {
      const int k = dev_kind<R>(i);
      for (size_t j = 0; j < h[i].size(); j++) {
        if (k == 0) h[i][j] = ((double *)dev[i])[j];
        else if (k == 1) h[i][j] = ((float *)dev[i])[j];
        else h[i][j] = ((int *)dev[i])[j];
      }
      if (fwrite(h[i].data(), sizeof(double), h[i].size(), out) != h[i].size()) return 3;
    }
  }
  for (void *p : dev) free(p);
  if (full == 2) fprintf(stderr, "fallbacks %lld\n", fallbacks);
  return 0;
}
}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: host_step IN OUT\n");
    return 1;
  }
  FILE *in = fopen(argv[1], "rb"), *out = fopen(argv[2], "wb");
  if (!in || !out) return 1;
  int64_t mp = 0;
  int32_t nsteps = 0, full = 0, precision = 0;
  double dels = 0, glob[18];
  if (!rd(in, &mp, 8) || !rd(in, &nsteps, 4) || !rd(in, &full, 4) || !rd(in, &precision, 4) || !rd(in, &dels, 8) ||
      !rd(in, glob, sizeof glob))
    return 2;
  std::vector<double> vtab(NVEGCLASS * VC_N), stab(NSOILCLASS * SC_N);
  if (!rd(in, vtab.data(), sizeof(double) * vtab.size()) || !rd(in, stab.data(), sizeof(double) * stab.size()))
    return 2;
  const int r = precision == 0 ? run<double>(in, out, mp, nsteps, full, dels, glob, vtab, stab)
                               : run<float>(in, out, mp, nsteps, full, dels, glob, vtab, stab);
  fclose(in);
  fclose(out);
  return r;
}

// cable_gpu.hip -- kernels and C ABI of libcable_gpu.so (include/cable_gpu.h).
//
// The context: device copies of the fields, resident forcing and class
// tables, the tile order, the stream and its timing events.  One launch of
// cbm_kernel (cbm_kernel.h) performs one timestep for every tile; HBM traffic
// per tile-step is the tile's state, parameters, forcing and outputs, read
// and written once.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>
#include <thread>
#include <atomic>
#include <algorithm>

#include "../../include/cable_gpu.h"
#include "cbm_kernel.h"

using namespace cbm;

// ---------------------------------------------------------------------------
// field table
namespace {
struct FieldMeta {
  const char *name;
  char prec;
  int n;
  int role;
};
#define PARAM ROLE_PARAM
#define FORC ROLE_FORC
#define STATE ROLE_STATE
#define DIAG ROLE_DIAG
const FieldMeta kFields[CABLE_NFIELDS] = {
#define X(id, name, prec, n, role) {name, #prec[0], n, role},
    CABLE_FIELDS(X)
#undef X
};
#undef PARAM
#undef FORC
#undef STATE
#undef DIAG

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIPCHECK(expr)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(CABLE_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// element size of a field on the device for a precision mode
size_t dev_elem(const FieldMeta &f, int precision) {
  switch (f.prec) {
    case 'I': return 4;
    case 'D': case 'B': return 8;
    default: return precision == CABLE_PREC_MIXED ? 4 : 8;
  }
}
}  // namespace

// ---------------------------------------------------------------------------
// The hybrid run's scratch for cable_gpu_run_hybrid_noalloc: the network's
// predictions of a step (NOUT floats per tile) and two counters, in a device
// array of the code object, so that the run allocates nothing.  Every call
// on a GPU shares that GPU's copy; calls that overlap on one GPU race.

namespace {
constexpr int64_t NOALLOC_MAX_TILES = 1048576;
constexpr size_t NOALLOC_PRED_BYTES = sizeof(float) * cbm_mlp::NOUT * (size_t)NOALLOC_MAX_TILES;
constexpr size_t NOALLOC_BYTES = NOALLOC_PRED_BYTES + 256;
constexpr int MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
// the device address of `scratch` on each device (the host-side name is a
// stand-in, not a device address)
void *g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// counters to zero with a kernel (a fill node, not a memset, under capture)
__global__ void zero_counters(unsigned long long *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// which of nst (at most 64) consecutive steps of the resident forcing window,
// the first at a.foff, have a leaf-bearing tile in daylight: bit s of *mask
// for step s, and of blocks[b] for the tiles of block b (256 tiles, the
// inference's block; zeroed by the caller).  The test is the inference's own
// (cbm::mlp_infer_kernel).
constexpr int LIT_CHUNK = 64;
__global__ void __launch_bounds__(256) lit_steps_kernel(cbm::StepArgs a, int nst, unsigned long long *mask,
                                                        unsigned long long *blocks) {
  const int *iveg = (const int *)a.F[CABLE_F_veg_iveg];
  const float *fsd = (const float *)a.F[CABLE_F_met_fsd] + a.foff;
  unsigned long long m = 0;
  // whole waves step together, so that a wave's 64 tiles lie in one block
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); t0 < a.mp; t0 += stride) {
    const int64_t t = t0 + (threadIdx.x & 63);
    unsigned long long mt = 0;
    if (t < a.mp && iveg[t] < 14)
      for (int s = 0; s < nst; s++) {
        const float f = fsd[(int64_t)s * a.mp + t];
        if (f + f > 0.001f) mt |= 1ull << s;
      }
    // an OR over the wave, one atomic per wave and block
    for (int k = 32; k > 0; k >>= 1) mt |= __shfl_xor(mt, k, 64);
    if ((threadIdx.x & 63) == 0 && mt) atomicOr(&blocks[t0 / 256], mt);
    m |= mt;
  }
  if ((threadIdx.x & 63) == 0 && m) atomicOr(mask, m);
}
}  // namespace

// ---------------------------------------------------------------------------
// context

struct cable_gpu_ctx {
  int64_t mp = 0;
  int precision = 0;
  std::vector<void *> dev;          // per field device arrays (n*mp elements)
  void **d_tab = nullptr;           // device copy of `dev`
  void **d_tab_run = nullptr;       // `dev` with forcing ids -> resident forcing
  std::vector<void *> forc;         // resident forcing arrays (8)
  int nsteps_forc = 0;
  int forc_k0 = 0;                  // step of the first resident forcing slot
  std::vector<double> forc_doy;     // per-step day of year of the window (set_met), or empty
  double ca = 400e-6, doy = 1.0;
  double zse[6] = {0}, zshh[7] = {0}, ratecp[3] = {0}, ratecs[2] = {0};
  double *d_vtab = nullptr, *d_stab = nullptr;
  bool have_tables = false;
  bool first_call = true;
  unsigned long long *d_npass = nullptr;
  int64_t last_passes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  // per-launch events of the hybrid run (a pair per step and kernel), and
  // the device time of its step and inference kernels
  std::vector<hipEvent_t> kev;
  double last_step_ms = 0.0, last_infer_ms = 0.0;
  bool launch_timing = true;         // record those events (cable_gpu_set_launch_timing)
  int device = 0;
  // device position j holds the caller's tile order[j] (empty: identity)
  std::vector<int64_t> order;
  bool touched = false;             // device fields hold anything but zeros
  // ---- surrogate (mixed contexts) ----
  float *hyb = nullptr;             // [3][mp] previous-step leaf/canopy-air state
  bool hyb_valid = false;
  float *sx = nullptr, *sy = nullptr;  // captured training samples
  unsigned long long *scount = nullptr;
  int64_t scap = 0;
  float *sparams = nullptr, *snorm = nullptr, *pred = nullptr;  // network, [mp][8]
  uint16_t *swt = nullptr;          // transposed BF16 weights, per layer [out][in]
  unsigned long long *nfallback = nullptr;  // tiles that fell back, summed over a run
  unsigned long long *d_lit = nullptr;      // the hybrid run's daylight steps (bit mask)
  unsigned long long *d_litblk = nullptr;   // the same per block of 256 tiles
  int64_t last_fallbacks = 0, last_tile_steps = 0;
  // ---- the driver's consistency check ----
  bool cons_on = false;
  double *cons_part = nullptr;      // per-workgroup sums of a step
  double *cons_acc = nullptr;       // count_bal, new_sumbal, new_sumfpn, new_sumfe
  // ---- batched transfers: a pinned host staging buffer and its device
  // twin ([transfer table | fields, each at a 256-byte offset]) ----
  unsigned char *h_stage = nullptr, *d_stage = nullptr;
  size_t stage_bytes = 0;
};

namespace {
// ---- batched transfers ----
// One entry per field of a batched transfer: the field's device array, its
// offset in the staging buffer and its length in 4-byte words.
struct XferEntry {
  void *dev;
  unsigned long long off, words;
};
constexpr size_t XFER_TABLE_BYTES = sizeof(XferEntry) * CABLE_NFIELDS;
constexpr size_t XFER_ALIGN = 256;

// stage -> fields (upload) or fields -> stage (download); blockIdx.y is the
// entry, the x blocks stride over its words
__global__ void xfer_kernel(const XferEntry *__restrict__ tab, uint32_t *__restrict__ stage, int to_fields) {
  const XferEntry e = tab[blockIdx.y];
  uint32_t *st = stage + e.off / 4, *dv = (uint32_t *)e.dev;
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < e.words;
       i += (unsigned long long)gridDim.x * blockDim.x) {
    if (to_fields)
      dv[i] = st[i];
    else
      st[i] = dv[i];
  }
}

// host values of one field <-> its device-typed image in the staging
// buffer, elements [i0, i1) (tile order and type conversion applied)
void pack_range(const cable_gpu_ctx *c, const FieldMeta &f, const void *host, int ht, unsigned char *img, size_t i0,
                size_t i1, bool to_device) {
  const size_t de = dev_elem(f, c->precision), mp = (size_t)c->mp;
  const int64_t *ord = c->order.empty() ? nullptr : c->order.data();
  const bool same = (f.prec == 'I' && ht == CABLE_HOST_I32) || (f.prec != 'I' && de == 8 && ht == CABLE_HOST_F64) ||
                    (f.prec != 'I' && de == 4 && ht == CABLE_HOST_F32);
  if (same && !ord) {
    if (to_device)
      memcpy(img + i0 * de, (const unsigned char *)host + i0 * de, (i1 - i0) * de);
    else
      memcpy((unsigned char *)host + i0 * de, img + i0 * de, (i1 - i0) * de);
    return;
  }
  for (size_t i = i0; i < i1; i++) {
    // rows of mp tiles; device position j of a row holds caller tile ord[j]
    const size_t h = ord ? (i - i % mp) + (size_t)ord[i % mp] : i;
    if (to_device) {
      const double v = ht == CABLE_HOST_F64 ? ((const double *)host)[h]
                       : ht == CABLE_HOST_F32 ? (double)((const float *)host)[h]
                                              : (double)((const int32_t *)host)[h];
      if (f.prec == 'I')
        ((int32_t *)img)[i] = (int32_t)llround(v);
      else if (de == 8)
        ((double *)img)[i] = v;
      else
        ((float *)img)[i] = (float)v;
    } else {
      const double v = f.prec == 'I' ? (double)((const int32_t *)img)[i]
                       : de == 8     ? ((const double *)img)[i]
                                     : (double)((const float *)img)[i];
      if (ht == CABLE_HOST_F64)
        ((double *)host)[h] = v;
      else if (ht == CABLE_HOST_F32)
        ((float *)host)[h] = (float)v;
      else
        ((int32_t *)host)[h] = (int32_t)llround(v);
    }
  }
}

int batched_transfer(cable_gpu_ctx *c, int n, const int *ids, void *const *hosts, const int *types, bool up);
}  // namespace

namespace {
size_t field_bytes(const cable_gpu_ctx *c, int id) {
  return dev_elem(kFields[id], c->precision) * (size_t)kFields[id].n * (size_t)c->mp;
}

// Several fields in one copy each way: the host side packs (unpacks) the
// fields' device-typed images in the pinned staging buffer on up to eight
// threads, one copy moves the buffer, one kernel scatters (gathers) the
// fields on the device.  The transfer table rides in front of the upload's
// data; a download sends it ahead in a copy of its own.
int batched_transfer(cable_gpu_ctx *c, int n, const int *ids, void *const *hosts, const int *types, bool up) {
  const char *who = up ? "cable_gpu_upload_fields" : "cable_gpu_download_fields";
  if (n > CABLE_NFIELDS) return fail(CABLE_E_ARG, "%s: %d fields (at most %d)", who, n, CABLE_NFIELDS);
  std::vector<XferEntry> tab(n);
  size_t off = XFER_TABLE_BYTES, maxw = 0;
  for (int k = 0; k < n; k++) {
    if (ids[k] < 0 || ids[k] >= CABLE_NFIELDS || !hosts[k])
      return fail(CABLE_E_ARG, "%s: bad field %d (entry %d)", who, ids[k], k);
    if (types[k] != CABLE_HOST_F64 && types[k] != CABLE_HOST_F32 && types[k] != CABLE_HOST_I32)
      return fail(CABLE_E_ARG, "%s: bad host type %d (entry %d)", who, types[k], k);
    const size_t b = field_bytes(c, ids[k]);
    tab[k] = {c->dev[ids[k]], off, b / 4};
    maxw = std::max(maxw, (size_t)(b / 4));
    off += (b + XFER_ALIGN - 1) / XFER_ALIGN * XFER_ALIGN;
  }
  if (n == 0) return CABLE_OK;
  if (off > c->stage_bytes) {
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (c->h_stage) HIPCHECK(hipHostFree(c->h_stage));
    if (c->d_stage) HIPCHECK(hipFree(c->d_stage));
    c->h_stage = c->d_stage = nullptr;
    c->stage_bytes = 0;
    HIPCHECK(hipHostMalloc((void **)&c->h_stage, off, hipHostMallocDefault));
    HIPCHECK(hipMalloc((void **)&c->d_stage, off));
    c->stage_bytes = off;
  }
  memcpy(c->h_stage, tab.data(), sizeof(XferEntry) * n);
  // the fields cut in pieces of at most 4 Mi elements, shared out to threads
  struct Piece {
    int k;
    size_t i0, i1;
  };
  std::vector<Piece> pieces;
  for (int k = 0; k < n; k++) {
    const size_t cnt = (size_t)kFields[ids[k]].n * (size_t)c->mp, step = (size_t)1 << 22;
    for (size_t i = 0; i < cnt; i += step) pieces.push_back({k, i, std::min(cnt, i + step)});
  }
  auto work = [&](bool to_device) {
    std::atomic<size_t> next(0);
    auto run = [&]() {
      for (size_t q; (q = next.fetch_add(1)) < pieces.size();) {
        const Piece &pc = pieces[q];
        pack_range(c, kFields[ids[pc.k]], hosts[pc.k], types[pc.k], c->h_stage + tab[pc.k].off, pc.i0, pc.i1,
                   to_device);
      }
    };
    const unsigned nt = (unsigned)std::min<size_t>({8, pieces.size(), std::max(1u, std::thread::hardware_concurrency())});
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) th.emplace_back(run);
    run();
    for (auto &t : th) t.join();
  };
  const dim3 grid((unsigned)std::min<size_t>(1024, (maxw + 255) / 256), (unsigned)n);
  if (up) {
    c->touched = true;
    work(true);
    HIPCHECK(hipMemcpyAsync(c->d_stage, c->h_stage, off, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(xfer_kernel, grid, dim3(256), 0, c->stream, (const XferEntry *)c->d_stage,
                       (uint32_t *)c->d_stage, 1);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(c->stream));
  } else {
    HIPCHECK(hipMemcpyAsync(c->d_stage, c->h_stage, sizeof(XferEntry) * n, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(xfer_kernel, grid, dim3(256), 0, c->stream, (const XferEntry *)c->d_stage,
                       (uint32_t *)c->d_stage, 0);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(c->h_stage + XFER_TABLE_BYTES, c->d_stage + XFER_TABLE_BYTES, off - XFER_TABLE_BYTES,
                            hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    work(false);
  }
  return CABLE_OK;
}

StepArgs make_args(const cable_gpu_ctx *c, void **tab, int ktau, double dels, int64_t foff) {
  StepArgs a;
  memset(&a, 0, sizeof a);
  a.F = tab;
  a.mp = c->mp;
  a.foff = foff;
  a.ktau = ktau;
  a.first_call = c->first_call ? 1 : 0;
  a.dels = dels;
  memcpy(a.zse, c->zse, sizeof a.zse);
  memcpy(a.zshh, c->zshh, sizeof a.zshh);
  memcpy(a.ratecp, c->ratecp, sizeof a.ratecp);
  memcpy(a.ratecs, c->ratecs, sizeof a.ratecs);
  a.vtab = c->d_vtab;
  a.stab = c->d_stab;
  a.ca = c->ca;
  a.doy = c->doy;
  if (tab == c->d_tab_run && !c->forc_doy.empty()) {
    // the window's day of year at this step (cable_gpu_set_met)
    const int64_t slot = (int64_t)ktau - 1 - c->forc_k0;
    if (slot >= 0 && slot < (int64_t)c->forc_doy.size()) a.doy = c->forc_doy[(size_t)slot];
  }
  a.npass = c->d_npass;
  a.driver = 1;
  return a;
}

hipError_t launch(const cable_gpu_ctx *c, const StepArgs &a, bool full) {
  return c->precision == CABLE_PREC_MIXED ? launch_mixed(full, a, c->mp, c->stream)
                                          : launch_fp64(full, a, c->mp, c->stream);
}

// host conversion between a host array and the device element type
// the host array has the device's element type and tile order: copy as is
bool direct_copy(const cable_gpu_ctx *c, const FieldMeta &f, int host_type) {
  if (!c->order.empty()) return false;
  if (f.prec == 'I') return host_type == CABLE_HOST_I32;
  const size_t de = dev_elem(f, c->precision);
  return (de == 8 && host_type == CABLE_HOST_F64) || (de == 4 && host_type == CABLE_HOST_F32);
}

int copy_to_device(cable_gpu_ctx *c, void *dst, const FieldMeta &f, const void *host, int host_type,
                   size_t count) {
  size_t de = dev_elem(f, c->precision);
  if (direct_copy(c, f, host_type)) {
    HIPCHECK(hipMemcpyAsync(dst, host, de * count, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CABLE_OK;
  }
  std::vector<unsigned char> buf(de * count);
  const size_t mp = (size_t)c->mp;
  const int64_t *ord = c->order.empty() ? nullptr : c->order.data();
  for (size_t i = 0; i < count; i++) {
    // rows of mp tiles; device position j of a row takes caller tile ord[j]
    const size_t s = ord ? (i - i % mp) + (size_t)ord[i % mp] : i;
    double v;
    switch (host_type) {
      case CABLE_HOST_F64: v = ((const double *)host)[s]; break;
      case CABLE_HOST_F32: v = ((const float *)host)[s]; break;
      case CABLE_HOST_I32: v = ((const int32_t *)host)[s]; break;
      default: return fail(CABLE_E_ARG, "bad host type %d", host_type);
    }
    if (f.prec == 'I')
      ((int32_t *)buf.data())[i] = (int32_t)llround(v);
    else if (de == 8)
      ((double *)buf.data())[i] = v;
    else
      ((float *)buf.data())[i] = (float)v;
  }
  HIPCHECK(hipMemcpyAsync(dst, buf.data(), buf.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return CABLE_OK;
}
}  // namespace

extern "C" {

int cable_gpu_nfields(void) { return CABLE_NFIELDS; }

int cable_gpu_field_info(int id, const char **name, char *prec, int *n, int *role) {
  if (id < 0 || id >= CABLE_NFIELDS) return fail(CABLE_E_ARG, "field id %d out of range", id);
  if (name) *name = kFields[id].name;
  if (prec) *prec = kFields[id].prec;
  if (n) *n = kFields[id].n;
  if (role) *role = kFields[id].role;
  return CABLE_OK;
}

int cable_gpu_field_id(const char *name) {
  if (!name) return -1;
  for (int i = 0; i < CABLE_NFIELDS; i++)
    if (strcmp(kFields[i].name, name) == 0) return i;
  return -1;
}

const char *cable_gpu_last_error(void) { return g_last_error.c_str(); }

int cable_gpu_create(cable_gpu_ctx **out, int64_t mp, int precision) {
  if (!out || mp <= 0) return fail(CABLE_E_ARG, "cable_gpu_create: bad arguments (mp=%lld)", (long long)mp);
  if (precision != CABLE_PREC_FP64 && precision != CABLE_PREC_MIXED)
    return fail(CABLE_E_ARG, "cable_gpu_create: unknown precision %d", precision);
  cable_gpu_ctx *c = new cable_gpu_ctx();
  c->mp = mp;
  c->precision = precision;
  c->dev.assign(CABLE_NFIELDS, nullptr);
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess && (c->device < 0 || c->device >= MAX_DEVICES)) {
    delete c;
    return fail(CABLE_E_ARG, "cable_gpu_create: device %d beyond the %d supported", c->device, MAX_DEVICES);
  }
  if (e == hipSuccess) {
    // the scratch array's address on this device, once per device, with the
    // device current
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[c->device]) {
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e == hipSuccess) g_scratch[c->device] = p;
    }
  }
  if (e == hipSuccess) e = hipStreamCreate(&c->stream);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  for (int i = 0; i < CABLE_NFIELDS && e == hipSuccess; i++) {
    e = hipMalloc(&c->dev[i], field_bytes(c, i));
    if (e == hipSuccess) e = hipMemsetAsync(c->dev[i], 0, field_bytes(c, i), c->stream);
  }
  if (e == hipSuccess) e = hipMalloc(&c->d_tab, sizeof(void *) * CABLE_NFIELDS);
  if (e == hipSuccess) e = hipMalloc(&c->d_tab_run, sizeof(void *) * CABLE_NFIELDS);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->d_tab, c->dev.data(), sizeof(void *) * CABLE_NFIELDS, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMalloc(&c->d_npass, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    cable_gpu_destroy(c);
    return fail(CABLE_E_HIP, "cable_gpu_create: %s", hipGetErrorString(e));
  }
  *out = c;
  return CABLE_OK;
}

int cable_gpu_destroy(cable_gpu_ctx *c) {
  if (!c) return CABLE_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  for (void *p : c->dev)
    if (p) hipFree(p);
  for (void *p : c->forc)
    if (p) hipFree(p);
  if (c->d_tab) hipFree(c->d_tab);
  if (c->d_tab_run) hipFree(c->d_tab_run);
  if (c->d_vtab) hipFree(c->d_vtab);
  if (c->d_stab) hipFree(c->d_stab);
  if (c->d_npass) hipFree(c->d_npass);
  void *sur[] = {c->cons_part, c->cons_acc, c->hyb, c->sx, c->sy, c->scount, c->sparams, c->snorm, c->pred, c->swt,
                 c->nfallback, c->d_lit, c->d_litblk};
  for (void *p : sur)
    if (p) hipFree(p);
  if (c->h_stage) hipHostFree(c->h_stage);
  if (c->d_stage) hipFree(c->d_stage);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  for (hipEvent_t ev : c->kev) hipEventDestroy(ev);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return CABLE_OK;
}

int cable_gpu_upload(cable_gpu_ctx *c, int id, const void *host, int host_type) {
  if (!c || !host || id < 0 || id >= CABLE_NFIELDS) return fail(CABLE_E_ARG, "cable_gpu_upload: bad arguments");
  c->touched = true;
  return copy_to_device(c, c->dev[id], kFields[id], host, host_type, (size_t)kFields[id].n * c->mp);
}

int cable_gpu_download(cable_gpu_ctx *c, int id, void *host, int host_type) {
  if (!c || !host || id < 0 || id >= CABLE_NFIELDS) return fail(CABLE_E_ARG, "cable_gpu_download: bad arguments");
  const FieldMeta &f = kFields[id];
  size_t count = (size_t)f.n * c->mp, de = dev_elem(f, c->precision);
  if (direct_copy(c, f, host_type)) {
    HIPCHECK(hipMemcpyAsync(host, c->dev[id], de * count, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CABLE_OK;
  }
  std::vector<unsigned char> buf(de * count);
  HIPCHECK(hipMemcpyAsync(buf.data(), c->dev[id], buf.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  const size_t mp = (size_t)c->mp;
  const int64_t *ord = c->order.empty() ? nullptr : c->order.data();
  for (size_t i = 0; i < count; i++) {
    const size_t d = ord ? (i - i % mp) + (size_t)ord[i % mp] : i;
    double v = f.prec == 'I' ? (double)((int32_t *)buf.data())[i]
               : de == 8     ? ((double *)buf.data())[i]
                             : (double)((float *)buf.data())[i];
    switch (host_type) {
      case CABLE_HOST_F64: ((double *)host)[d] = v; break;
      case CABLE_HOST_F32: ((float *)host)[d] = (float)v; break;
      case CABLE_HOST_I32: ((int32_t *)host)[d] = (int32_t)llround(v); break;
      default: return fail(CABLE_E_ARG, "bad host type %d", host_type);
    }
  }
  return CABLE_OK;
}

int cable_gpu_upload_fields(cable_gpu_ctx *c, int n, const int *ids, const void *const *hosts,
                            const int *host_types) {
  if (!c || n < 0 || (n && (!ids || !hosts || !host_types)))
    return fail(CABLE_E_ARG, "cable_gpu_upload_fields: bad arguments");
  return batched_transfer(c, n, ids, const_cast<void *const *>(hosts), host_types, true);
}

int cable_gpu_download_fields(cable_gpu_ctx *c, int n, const int *ids, void *const *hosts, const int *host_types) {
  if (!c || n < 0 || (n && (!ids || !hosts || !host_types)))
    return fail(CABLE_E_ARG, "cable_gpu_download_fields: bad arguments");
  return batched_transfer(c, n, ids, hosts, host_types, false);
}

// ---------------------------------------------------------------------------
// state save and load (restart)

}  // extern "C"

namespace {
// header of a saved state: magic, layout version, mp, precision, field count
// and the field table's checksum (32 bytes); then per saved field a 16-byte
// record (id, n, element size, 0) and the field's device-typed values in the
// caller's tile order, padded to a multiple of 8 bytes, so that every
// record and every array starts 8-byte aligned
constexpr uint32_t STATE_MAGIC = 0x534C4243u;  // "CBLS"
constexpr uint32_t STATE_VERSION = 2;
constexpr int64_t STATE_REC_BYTES = 4 * sizeof(int32_t);
int64_t pad8(int64_t b) { return (b + 7) & ~(int64_t)7; }
struct StateHeader {
  uint32_t magic, version;
  int64_t mp;
  int32_t precision, nsaved;
  uint64_t table_sum;
};
static_assert(sizeof(StateHeader) == 32, "the saved state's header is 32 bytes");
bool saved_field(int id) { return kFields[id].role == ROLE_STATE || kFields[id].role == ROLE_DIAG; }
uint64_t table_checksum() {
  uint64_t h = 1469598103934665603ull;  // FNV-1a over names, precisions and counts
  for (int id = 0; id < CABLE_NFIELDS; id++) {
    for (const char *q = kFields[id].name; *q; q++) h = (h ^ (uint8_t)*q) * 1099511628211ull;
    h = (h ^ (uint8_t)kFields[id].prec) * 1099511628211ull;
    h = (h ^ (uint64_t)kFields[id].n) * 1099511628211ull;
  }
  return h;
}
int64_t state_bytes(const cable_gpu_ctx *c) {
  int64_t b = sizeof(StateHeader);
  for (int id = 0; id < CABLE_NFIELDS; id++)
    if (saved_field(id)) b += STATE_REC_BYTES + pad8((int64_t)field_bytes(c, id));
  return b;
}
int host_type_of(const FieldMeta &f, int precision) {
  if (f.prec == 'I') return CABLE_HOST_I32;
  return dev_elem(f, precision) == 8 ? CABLE_HOST_F64 : CABLE_HOST_F32;
}
}  // namespace

extern "C" {

int64_t cable_gpu_state_bytes(cable_gpu_ctx *c) {
  if (!c) return -(int64_t)fail(CABLE_E_ARG, "cable_gpu_state_bytes: null context");
  return state_bytes(c);
}

int cable_gpu_save_state(cable_gpu_ctx *c, void *buf, int64_t bytes) {
  if (!c || !buf) return fail(CABLE_E_ARG, "cable_gpu_save_state: null argument");
  if (bytes < state_bytes(c))
    return fail(CABLE_E_ARG, "cable_gpu_save_state: buffer of %lld bytes, %lld needed", (long long)bytes,
                (long long)state_bytes(c));
  unsigned char *p = (unsigned char *)buf;
  StateHeader h = {STATE_MAGIC, STATE_VERSION, c->mp, c->precision, 0, table_checksum()};
  for (int id = 0; id < CABLE_NFIELDS; id++) h.nsaved += saved_field(id) ? 1 : 0;
  memcpy(p, &h, sizeof h);
  p += sizeof h;
  for (int id = 0; id < CABLE_NFIELDS; id++) {
    if (!saved_field(id)) continue;
    const FieldMeta &f = kFields[id];
    const int32_t rec[4] = {id, f.n, (int32_t)dev_elem(f, c->precision), 0};
    memcpy(p, rec, sizeof rec);
    p += sizeof rec;
    int r = cable_gpu_download(c, id, p, host_type_of(f, c->precision));
    if (r) return r;
    const int64_t fb = (int64_t)field_bytes(c, id);
    memset(p + fb, 0, (size_t)(pad8(fb) - fb));
    p += pad8(fb);
  }
  return CABLE_OK;
}

int cable_gpu_load_state(cable_gpu_ctx *c, const void *buf, int64_t bytes, int continue_run) {
  if (!c || !buf) return fail(CABLE_E_ARG, "cable_gpu_load_state: null argument");
  if (bytes < (int64_t)sizeof(StateHeader)) return fail(CABLE_E_ARG, "cable_gpu_load_state: truncated state");
  const unsigned char *p = (const unsigned char *)buf;
  StateHeader h;
  memcpy(&h, p, sizeof h);
  if (h.magic != STATE_MAGIC || h.version != STATE_VERSION)
    return fail(CABLE_E_ARG, "cable_gpu_load_state: not a saved state of this library");
  if (h.table_sum != table_checksum())
    return fail(CABLE_E_ARG, "cable_gpu_load_state: saved with a different field table");
  if (h.mp != c->mp || h.precision != c->precision)
    return fail(CABLE_E_ARG, "cable_gpu_load_state: saved for mp=%lld precision %d, context has mp=%lld precision %d",
                (long long)h.mp, h.precision, (long long)c->mp, c->precision);
  if (bytes < state_bytes(c)) return fail(CABLE_E_ARG, "cable_gpu_load_state: truncated state");
  p += sizeof h;
  for (int i = 0; i < h.nsaved; i++) {
    int32_t rec[4];
    memcpy(rec, p, sizeof rec);
    p += sizeof rec;
    const int id = rec[0];
    if (id < 0 || id >= CABLE_NFIELDS || !saved_field(id) || rec[1] != kFields[id].n ||
        rec[2] != (int32_t)dev_elem(kFields[id], c->precision))
      return fail(CABLE_E_ARG, "cable_gpu_load_state: bad field record %d", i);
    int r = cable_gpu_upload(c, id, p, host_type_of(kFields[id], c->precision));
    if (r) return r;
    p += pad8((int64_t)field_bytes(c, id));
  }
  // an exact continuation keeps the run's first-call state; otherwise the next
  // step is the first of a new run, as a reference run started from a restart
  c->first_call = continue_run == 0;
  c->hyb_valid = false;
  return CABLE_OK;
}

int cable_gpu_set_tile_order(cable_gpu_ctx *c, const int64_t *order) {
  if (!c || !order) return fail(CABLE_E_ARG, "cable_gpu_set_tile_order: null argument");
  if (c->nsteps_forc)
    return fail(CABLE_E_STATE, "cable_gpu_set_tile_order: call it before cable_gpu_set_forcing");
  std::vector<char> seen((size_t)c->mp, 0);
  for (int64_t j = 0; j < c->mp; j++) {
    if (order[j] < 0 || order[j] >= c->mp || seen[(size_t)order[j]])
      return fail(CABLE_E_ARG, "cable_gpu_set_tile_order: not a permutation of 0..%lld (entry %lld)",
                  (long long)(c->mp - 1), (long long)j);
    seen[(size_t)order[j]] = 1;
  }
  const std::vector<int64_t> old_order = c->order, new_order(order, order + c->mp);
  if (!c->touched) {  // still zero-filled: nothing to carry over
    c->order = new_order;
    return CABLE_OK;
  }
  // carry the device contents over to the new order (exact: every device
  // element type converts to double and back without change)
  std::vector<double> h;
  for (int id = 0; id < CABLE_NFIELDS; id++) {
    h.resize((size_t)kFields[id].n * (size_t)c->mp);
    c->order = old_order;
    int r = cable_gpu_download(c, id, h.data(), CABLE_HOST_F64);
    c->order = new_order;
    if (!r) r = cable_gpu_upload(c, id, h.data(), CABLE_HOST_F64);
    if (r) return r;
  }
  return CABLE_OK;
}

int cable_gpu_set_globals(cable_gpu_ctx *c, const double *zse, const double *zshh, const double *ratecp,
                          const double *ratecs) {
  if (!c || !zse || !zshh || !ratecp || !ratecs) return fail(CABLE_E_ARG, "cable_gpu_set_globals: null argument");
  memcpy(c->zse, zse, sizeof c->zse);
  memcpy(c->zshh, zshh, sizeof c->zshh);
  memcpy(c->ratecp, ratecp, sizeof c->ratecp);
  memcpy(c->ratecs, ratecs, sizeof c->ratecs);
  return CABLE_OK;
}

namespace {
// one full-diagnostic step: driver = 1 for the offline driver's step, 0 for
// cbm alone
int step_full(cable_gpu_ctx *c, int ktau, double dels, int driver, const char *who) {
  if (!c) return fail(CABLE_E_ARG, "%s: null context", who);
  if (ktau < 1 || !(dels > 0)) return fail(CABLE_E_ARG, "%s: bad ktau %d or dels %g", who, ktau, dels);
  HIPCHECK(hipMemsetAsync(c->d_npass, 0, sizeof(unsigned long long), c->stream));
  StepArgs a = make_args(c, c->d_tab, ktau, dels, 0);
  a.driver = driver;
  c->touched = true;
  hipError_t e = launch(c, a, true);
  if (e != hipSuccess) return fail(CABLE_E_HIP, "%s launch: %s", who, hipGetErrorString(e));
  unsigned long long np = 0;
  HIPCHECK(hipMemcpyAsync(&np, c->d_npass, sizeof np, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->last_passes = (int64_t)np;
  c->first_call = false;
  return CABLE_OK;
}
}  // namespace

int cable_gpu_step(cable_gpu_ctx *c, int ktau, double dels) {
  return step_full(c, ktau, dels, 1, "cable_gpu_step");
}

int cable_gpu_cbm(cable_gpu_ctx *c, int ktau, double dels) {
  return step_full(c, ktau, dels, 0, "cable_gpu_cbm");
}

int cable_gpu_veg_cols(void) { return VC_N; }
int cable_gpu_soil_cols(void) { return SC_N; }
int cable_gpu_table_column(int table, int id) {
  if (id < 0 || id >= CABLE_NFIELDS) return -1;
  switch (id) {
#define X(fid, name, prec, n, role) \
  case CABLE_F_##fid: return table == 0 ? veg_col(CABLE_F_##fid) : soil_col(CABLE_F_##fid);
    CABLE_FIELDS(X)
#undef X
  }
  return -1;
}

int cable_gpu_set_class_tables(cable_gpu_ctx *c, const double *veg, const double *soil) {
  if (!c || !veg || !soil) return fail(CABLE_E_ARG, "cable_gpu_set_class_tables: null argument");
  if (!c->d_vtab) HIPCHECK(hipMalloc(&c->d_vtab, sizeof(double) * NVEGCLASS * VC_N));
  if (!c->d_stab) HIPCHECK(hipMalloc(&c->d_stab, sizeof(double) * NSOILCLASS * SC_N));
  HIPCHECK(hipMemcpyAsync(c->d_vtab, veg, sizeof(double) * NVEGCLASS * VC_N, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->d_stab, soil, sizeof(double) * NSOILCLASS * SC_N, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->have_tables = true;
  return CABLE_OK;
}

int cable_gpu_set_forcing(cable_gpu_ctx *c, int nsteps, const void *fsd1, const void *fld, const void *tk,
                          const void *qv, const void *ua, const void *pmb, const void *precip, const void *coszen,
                          int host_type, double ca, double doy) {
  if (!c || nsteps <= 0) return fail(CABLE_E_ARG, "cable_gpu_set_forcing: bad arguments");
  const void *src[8] = {fsd1, fld, tk, qv, ua, pmb, precip, coszen};
  const int ids[8] = {CABLE_F_met_fsd, CABLE_F_met_fld, CABLE_F_met_tk, CABLE_F_met_qv,
                      CABLE_F_met_ua, CABLE_F_met_pmb, CABLE_F_met_precip, CABLE_F_met_coszen};
  for (int i = 0; i < 8; i++)
    if (!src[i]) return fail(CABLE_E_ARG, "cable_gpu_set_forcing: null forcing array %d", i);
  for (void *p : c->forc)
    if (p) hipFree(p);
  c->forc.assign(8, nullptr);
  FieldMeta fm = {"forcing", 'F', 1, ROLE_FORC};
  size_t count = (size_t)nsteps * c->mp;
  size_t bytes = dev_elem(fm, c->precision) * count;
  std::vector<void *> tab = c->dev;
  for (int i = 0; i < 8; i++) {
    HIPCHECK(hipMalloc(&c->forc[i], bytes));
    int r = copy_to_device(c, c->forc[i], fm, src[i], host_type, count);
    if (r) return r;
    tab[ids[i]] = c->forc[i];
  }
  HIPCHECK(hipMemcpyAsync(c->d_tab_run, tab.data(), sizeof(void *) * CABLE_NFIELDS, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->nsteps_forc = nsteps;
  c->forc_k0 = 0;
  c->forc_doy.clear();
  c->ca = ca;
  c->doy = doy;
  return CABLE_OK;
}

// ---------------------------------------------------------------------------
// synthetic forcing generated on the device

}  // extern "C"

namespace {
// ---- the offline driver's consistency check -------------------------------
constexpr int CONS_THREADS = 256;

// in-order tree over the threads of one workgroup (CONS_THREADS, a power of 2)
__device__ inline void block_sum(double (*sh)[4], int nq) {
  for (int w = CONS_THREADS / 2; w > 0; w >>= 1) {
    __syncthreads();
    if ((int)threadIdx.x < w)
      for (int q = 0; q < nq; q++) sh[threadIdx.x][q] += sh[threadIdx.x + w][q];
  }
  __syncthreads();
}

// one step's tile sums from the step kernel's per-workgroup sums, added to the
// running sums as src/offline/cable_serial.F90:748-752 does:
// new_sumbal += SUM(wbal)/mp + SUM(ebal)/mp, new_sumfpn += SUM(fpn)/mp,
// new_sumfe += SUM(fe)/mp, count_bal += 1
__global__ void __launch_bounds__(CONS_THREADS) cons_reduce_kernel(const double *part, int64_t nb, double mp,
                                                                  double *acc) {
  __shared__ double sh[CONS_THREADS][4];
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = threadIdx.x; i < nb; i += CONS_THREADS)
    for (int q = 0; q < 4; q++) s[q] += part[i * 4 + q];
  for (int q = 0; q < 4; q++) sh[threadIdx.x][q] = s[q];
  block_sum(sh, 4);
  if (threadIdx.x == 0) {
    acc[0] = acc[0] + 1.0;
    acc[1] = acc[1] + sh[0][0] / mp + sh[0][1] / mp;
    acc[2] = acc[2] + sh[0][2] / mp;
    acc[3] = acc[3] + sh[0][3] / mp;
  }
}

// SUM(x) over n doubles, one workgroup, fixed order
__global__ void __launch_bounds__(CONS_THREADS) sum_kernel(const double *x, int64_t n, double *out) {
  __shared__ double sh[CONS_THREADS][4];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += CONS_THREADS) s += x[i];
  sh[threadIdx.x][0] = s;
  block_sum(sh, 1);
  if (threadIdx.x == 0) *out = sh[0][0];
}

// workgroups of a step launch (launch_cbm: 64 threads for FP64, 256 for FP32)
int64_t step_workgroups(const cable_gpu_ctx *c) {
  const int64_t block = c->precision == CABLE_PREC_MIXED ? 256 : 64;
  return (c->mp + block - 1) / block;
}

hipError_t launch_cons_reduce(cable_gpu_ctx *c) {
  hipLaunchKernelGGL(cons_reduce_kernel, dim3(1), dim3(CONS_THREADS), 0, c->stream, c->cons_part,
                     step_workgroups(c), (double)c->mp, c->cons_acc);
  return hipGetLastError();
}

// ---- met forcing in the met file's variables and units ---------------------
// the reader's saturation vapour pressure and specific humidity from
// relative humidity, src/offline/cable_checks.F90:633-696, in precision R
template <class R>
__device__ inline R met_svp(R tk) {
  const R temp = tk - R(273.15);
  if (temp < R(-20.0)) {
    const R toot = R(273.15) / tk, toto = R(1.) / toot;
    const R eilog = R(-9.09718) * (toot - R(1)) - R(3.56654) * (log(toot) / log(R(10.0))) +
                    R(0.876793) * (R(1) - toto) + (log(R(6.1071)) / log(R(10.0)));
    return pow(R(10.0), eilog);
  }
  const R tsot = R(373.15) / tk;
  const R ewlog = R(-7.90298) * (tsot - R(1)) + R(5.02808) * (log(tsot) / log(R(10.0)));
  const R ewlog2 = ewlog - R(1.3816e-07) * (pow(R(10), R(11.344) * (R(1) - (R(1) / tsot))) - R(1));
  const R ewlog3 = ewlog2 + R(0.0081328) * (pow(R(10), R(-3.49149) * (tsot - R(1))) - R(1));
  const R ewlog4 = ewlog3 + (log(R(1013.246)) / log(R(10.0)));
  return pow(R(10.0), ewlog4);
}

struct MetArgs {
  const float *swdown, *lwdown, *tair, *qair, *wind, *wind_e, *psurf, *rainf;
  const int64_t *tix;       // device position -> caller tile
  const double *elev, *lat; // per caller tile
  const double *hod, *doy;  // per step of the window
  double tair_offset, psurf_scale, rainf_scale;
  int qair_is_rh;
};

// one window of forcing from the met variables, as get_met_data forms met%
// (src/offline/cable_input.F90): fsd(:,1) = fsd(:,2) = SWdown/2 (:1880-1883),
// tk = Tair + convert%Tair (:1923), pmb = PSurf * convert%PSurf (:1959) or,
// without PSurf, from elevation and tk (:1979-1981), qv = Qair or rh_sh of
// it (:2000-2003), ua = Wind or sqrt(Wind_N**2 + Wind_E**2) (:2094),
// precip = Rainf * convert%Rainf (:2202-2204), fld = LWdown or Swinbank's
// 0.0000094 * 0.0000000567 * tk**6 (:2257), coszen = sinbet(doy, lat, hod)
// (:2677, src/science/radiation/cbl_sinbet.F90:8-28)
template <class R>
__global__ void __launch_bounds__(256) met_kernel(int64_t mp, MetArgs m, R *fsd1, R *fld, R *tk, R *qv, R *ua,
                                                  R *pmb, R *precip, R *coszen) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= mp) return;
  const int64_t t = m.tix[j], k = blockIdx.y, src = k * mp + t, o = k * mp + j;
  const R tair = (R)m.tair[src] + (R)m.tair_offset;
  R ps;
  if (m.psurf) {
    ps = (R)m.psurf[src] * (R)m.psurf_scale;
  } else {
    ps = R(1013.25) * pow(tair / (tair + R(0.0065) * (R)m.elev[t]), R(9.80665) / R(287.04) / R(0.0065));
  }
  R q = (R)m.qair[src];
  if (m.qair_is_rh) {
    const R es = met_svp(tair);
    const R ws = R(0.622) * es / (ps - es);
    q = (q / R(100.0)) * ws;
  }
  R u = (R)m.wind[src];
  if (m.wind_e) {
    const R e = (R)m.wind_e[src];
    u = sqrt(u * u + e * e);
  }
  const R lw = m.lwdown ? (R)m.lwdown[src] : R(0.0000094) * R(0.0000000567) * pow(tair, R(6.0));
  const R pi = R(3.1415927), pi180 = pi / R(180.0);
  const R doy = (R)m.doy[k], hod = (R)m.hod[k], xslat = (R)m.lat[t];
  const R sindec = -sin(R(23.45) * pi180) * cos(R(2.) * pi * (doy + R(10.0)) / R(365.0));
  R z = sin(pi180 * xslat) * sindec + cos(pi180 * xslat) * sqrt(R(1.) - sindec * sindec) * cos(pi * (hod - R(12.0)) / R(12.0));
  z = z > R(1e-8) ? z : R(1e-8);
  fsd1[o] = R(0.5) * (R)m.swdown[src];
  tk[o] = tair;
  pmb[o] = ps;
  qv[o] = q;
  ua[o] = u;
  precip[o] = (R)m.rainf[src] * (R)m.rainf_scale;
  fld[o] = lw;
  coszen[o] = z;
}

// the reader's clock (src/offline/cable_input.F90:1685-1850, local time):
// hod = shod at the run's first step, then hod + dels/3600 per step; a day
// rolls over at 24 h with the month and year as the reader sets them
// (IS_LEAPYEAR, src/util/cable_common.F90:226-235)
template <class R>
struct MetClock {
  R hod, doy;
  int moy, year;
  static bool leap(int y) { return ((y % 4 == 0 && y % 100 != 0) || y % 400 == 0); }
  void roll() {
    if (hod < R(0.0)) {
      doy = doy - 1;
      hod = hod + R(24.0);
      static const int ml[] = {31, 60, 91, 121, 152, 182, 213, 244, 274, 305, 335};
      static const int mn[] = {31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334};
      const int *tab = leap(year) ? ml : mn;
      const int d = (int)doy;
      if (d == 0) {
        moy = 12;
        year = year - 1;
        // the reader assumes the year before a leap year is not one
        doy = leap(year + 1) ? R(365) : (leap(year) ? R(366) : R(365));
      }
      for (int i = 0; i < 11; i++)
        if (d == tab[i]) moy = i + 1;
    } else if (hod >= R(24.0)) {
      doy = doy + 1;
      hod = hod - R(24.0);
      static const int ml[] = {32, 61, 92, 122, 153, 183, 214, 245, 275, 306, 336};
      static const int mn[] = {32, 60, 91, 121, 152, 182, 213, 244, 274, 305, 335};
      const bool lp = leap(year);
      const int *tab = lp ? ml : mn;
      const int d = (int)doy;
      for (int i = 0; i < 11; i++)
        if (d == tab[i]) moy = i + 2;
      if (d == (lp ? 367 : 366)) {
        year = year + 1;
        moy = 1;
        doy = R(1);
      }
    }
  }
};

template <class R>
void met_clock(double shod, int sdoy, int smoy, int syear, double dels, int k_first, int nsteps,
               std::vector<double> &hod, std::vector<double> &doy) {
  MetClock<R> c{(R)shod, (R)sdoy, smoy, syear};
  c.roll();
  for (int k = 0; k < k_first + nsteps; k++) {
    if (k > 0) {
      c.hod = c.hod + (R)dels / R(3600.0);
      c.roll();
    }
    if (k >= k_first) {
      hod.push_back((double)c.hod);
      doy.push_back((double)c.doy);
    }
  }
}

// SplitMix64's increment and finaliser (Steele, Lea and Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014): mix64(0) is the
// first output of SplitMix64 seeded with 0, 0xE220A8397B1DCDAF
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

constexpr int NSYNTH = 8;  // draws per tile-step: sw, lw, rh, ua (2), rain occurrence, amount, psurf
struct SynthKeys {
  uint64_t k[NSYNTH];
};

// U(0,1), open at both ends, of draw `s` for (caller tile, step)
__device__ inline double synth_u(const SynthKeys &keys, int s, int64_t tile, int64_t step) {
  const uint64_t h = mix64(keys.k[s] ^ ((uint64_t)step << 32) ^ (uint64_t)tile);
  return ((double)(h >> 11) + 0.5) * 0x1.0p-53;
}

// saturation vapour pressure (hPa) and specific humidity from relative
// humidity, src/offline/cable_checks.F90:633-696
__device__ inline double synth_svp(double tk) {
  const double l10 = log(10.0);
  if (tk - 273.15 < -20.0) {
    const double toot = 273.15 / tk, toto = 1.0 / toot;
    const double eilog = -9.09718 * (toot - 1) - 3.56654 * (log(toot) / l10) + 0.876793 * (1 - toto) +
                         (log(6.1071) / l10);
    return pow(10.0, eilog);
  }
  const double tsot = 373.15 / tk;
  double ew = -7.90298 * (tsot - 1) + 5.02808 * (log(tsot) / l10);
  ew = ew - 1.3816e-07 * (pow(10.0, 11.344 * (1 - (1 / tsot))) - 1);
  ew = ew + 0.0081328 * (pow(10.0, -3.49149 * (tsot - 1)) - 1);
  ew = ew + (log(1013.246) / l10);
  return pow(10.0, ew);
}

// one step of the synthetic met of every tile, in the forcing reader's
// variables: SWdown from the diurnal cycle, split evenly between the two
// bands (src/offline/cable_input.F90:1880-1883); Tair from the tile's mean
// and a +-5 K diurnal swing; qv from RH ~ U(30, 90)% by rh_sh; lognormal wind
// (mean 3 m/s, sigma 0.5, Box-Muller); Bernoulli(0.1) x Exp(mean 2 mm/h)
// precipitation in mm per step; PSurf ~ U(950, 1020) hPa; LWdown ~
// U(250, 420) W/m2; coszen by sinbet (src/science/radiation/cbl_sinbet.F90:8-28)
template <class R>
__global__ void __launch_bounds__(256)
    synth_kernel(int64_t mp, int k_first, SynthKeys keys, double dels, double doy, double hod0,
                 const int64_t *__restrict__ tix, const double *__restrict__ t0, const double *__restrict__ lat,
                 R *fsd1, R *fld, R *tk, R *qv, R *ua, R *pmb, R *precip, R *coszen) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= mp) return;
  const int64_t k = k_first + (int64_t)blockIdx.y, o = (int64_t)blockIdx.y * mp + j, t = tix[j];
  const double two_pi = 2.0 * M_PI;
  const double hod = hod0 + (double)k * (dels / 3600.0);
  const double diurnal = fmax(0.0, 1000.0 * sin(two_pi * (hod - 6.0) / 24.0));
  const double sw = diurnal * (0.5 + 0.5 * synth_u(keys, 0, t, k));
  const double tair = t0[j] + 5.0 * sin(two_pi * (hod - 9.0) / 24.0);
  const double ps = 950.0 + 70.0 * synth_u(keys, 7, t, k);
  const double es = synth_svp(tair);
  const double rh = 100.0 * (0.3 + 0.6 * synth_u(keys, 2, t, k));
  const double q = (rh / 100.0) * (0.622 * es / (ps - es));
  const double z = sqrt(-2.0 * log(synth_u(keys, 3, t, k))) * cos(two_pi * synth_u(keys, 4, t, k));
  const double sigma = 0.5;
  const double wind = exp((log(3.0) - 0.5 * sigma * sigma) + sigma * z);
  const double rain = synth_u(keys, 5, t, k) < 0.1 ? 2.0 * -log(synth_u(keys, 6, t, k)) * (dels / 3600.0) : 0.0;
  // sinbet with the reference's pi = 3.1415927 (src/params/cable_maths_constants_mod.F90:32)
  const double pi_f = 3.1415927, d2r = pi_f / 180.0;
  const double sindec = -sin(23.45 * d2r) * cos(2.0 * pi_f * (doy + 10.0) / 365.0);
  const double cz = sin(d2r * lat[j]) * sindec +
                    cos(d2r * lat[j]) * sqrt(1.0 - sindec * sindec) * cos(pi_f * (hod - 12.0) / 12.0);
  fsd1[o] = (R)(0.5 * sw);
  fld[o] = (R)(250.0 + 170.0 * synth_u(keys, 1, t, k));
  tk[o] = (R)tair;
  qv[o] = (R)q;
  ua[o] = (R)wind;
  pmb[o] = (R)ps;
  precip[o] = (R)rain;
  coszen[o] = (R)fmax(cz, 1e-8);
}
}  // namespace

extern "C" {

int cable_gpu_synth_forcing(cable_gpu_ctx *c, int k_first, int nsteps, uint64_t seed, double dels, double doy,
                            double hod0, const double *t0, const double *lat) {
  if (!c || !t0 || !lat || k_first < 0 || nsteps <= 0 || nsteps > 65535 || !(dels > 0))
    return fail(CABLE_E_ARG, "cable_gpu_synth_forcing: bad arguments");
  const size_t mp = (size_t)c->mp;
  FieldMeta fm = {"forcing", 'F', 1, ROLE_FORC};
  const size_t bytes = dev_elem(fm, c->precision) * (size_t)nsteps * mp;
  if (c->nsteps_forc != nsteps || c->forc.size() != 8) {
    for (void *p : c->forc)
      if (p) HIPCHECK(hipFree(p));
    c->forc.assign(8, nullptr);
    c->nsteps_forc = 0;
    for (int i = 0; i < 8; i++) HIPCHECK(hipMalloc(&c->forc[i], bytes));
  }
  // per-tile inputs in device order; the draws are keyed by the caller's tile
  std::vector<int64_t> tix(mp);
  std::vector<double> ht0(mp), hlat(mp);
  for (size_t j = 0; j < mp; j++) {
    tix[j] = c->order.empty() ? (int64_t)j : c->order[j];
    ht0[j] = t0[tix[j]];
    hlat[j] = lat[tix[j]];
  }
  int64_t *d_tix = nullptr;
  double *d_t0 = nullptr, *d_lat = nullptr;
  HIPCHECK(hipMalloc(&d_tix, sizeof(int64_t) * mp));
  HIPCHECK(hipMalloc(&d_t0, sizeof(double) * mp));
  HIPCHECK(hipMalloc(&d_lat, sizeof(double) * mp));
  HIPCHECK(hipMemcpyAsync(d_tix, tix.data(), sizeof(int64_t) * mp, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(d_t0, ht0.data(), sizeof(double) * mp, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(d_lat, hlat.data(), sizeof(double) * mp, hipMemcpyHostToDevice, c->stream));
  SynthKeys keys;
  for (int s = 0; s < NSYNTH; s++) keys.k[s] = mix64(mix64(seed) ^ (uint64_t)(s + 1));
  const dim3 grid((unsigned)((mp + 255) / 256), (unsigned)nsteps);
  void **f = c->forc.data();
  if (c->precision == CABLE_PREC_MIXED)
    hipLaunchKernelGGL(synth_kernel<float>, grid, dim3(256), 0, c->stream, c->mp, k_first, keys, dels, doy, hod0,
                       d_tix, d_t0, d_lat, (float *)f[0], (float *)f[1], (float *)f[2], (float *)f[3],
                       (float *)f[4], (float *)f[5], (float *)f[6], (float *)f[7]);
  else
    hipLaunchKernelGGL(synth_kernel<double>, grid, dim3(256), 0, c->stream, c->mp, k_first, keys, dels, doy, hod0,
                       d_tix, d_t0, d_lat, (double *)f[0], (double *)f[1], (double *)f[2], (double *)f[3],
                       (double *)f[4], (double *)f[5], (double *)f[6], (double *)f[7]);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipFree(d_tix));
  HIPCHECK(hipFree(d_t0));
  HIPCHECK(hipFree(d_lat));
  const int ids[8] = {CABLE_F_met_fsd, CABLE_F_met_fld, CABLE_F_met_tk, CABLE_F_met_qv,
                      CABLE_F_met_ua, CABLE_F_met_pmb, CABLE_F_met_precip, CABLE_F_met_coszen};
  std::vector<void *> tab = c->dev;
  for (int i = 0; i < 8; i++) tab[ids[i]] = c->forc[i];
  HIPCHECK(hipMemcpyAsync(c->d_tab_run, tab.data(), sizeof(void *) * CABLE_NFIELDS, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->nsteps_forc = nsteps;
  c->forc_k0 = k_first;
  c->forc_doy.clear();
  c->ca = 400e-6;
  c->doy = doy;
  return CABLE_OK;
}

int cable_gpu_set_met(cable_gpu_ctx *c, int k_first, int nsteps, double dels, const cable_gpu_met *m) {
  if (!c || !m || k_first < 0 || nsteps <= 0 || nsteps > 65535 || !(dels > 0))
    return fail(CABLE_E_ARG, "cable_gpu_set_met: bad arguments");
  if (!m->swdown || !m->tair || !m->qair || !m->wind || !m->rainf || !m->latitude || (!m->psurf && !m->elevation))
    return fail(CABLE_E_ARG, "cable_gpu_set_met: SWdown, Tair, Qair, Wind, Rainf and latitude are required, "
                             "and elevation when PSurf is absent");
  const size_t mp = (size_t)c->mp, n = (size_t)nsteps * mp;
  // the clock of the window, uniform over the tiles (local time)
  std::vector<double> hod, doy;
  if (c->precision == CABLE_PREC_MIXED)
    met_clock<float>(m->shod, m->sdoy, m->smoy, m->syear, dels, k_first, nsteps, hod, doy);
  else
    met_clock<double>(m->shod, m->sdoy, m->smoy, m->syear, dels, k_first, nsteps, hod, doy);
  FieldMeta fm = {"forcing", 'F', 1, ROLE_FORC};
  const size_t bytes = dev_elem(fm, c->precision) * n;
  if (c->nsteps_forc != nsteps || c->forc.size() != 8) {
    for (void *p : c->forc)
      if (p) HIPCHECK(hipFree(p));
    c->forc.assign(8, nullptr);
    c->nsteps_forc = 0;
    for (int i = 0; i < 8; i++) HIPCHECK(hipMalloc(&c->forc[i], bytes));
  }
  // inputs to the device: the met variables in the caller's tile order
  std::vector<void *> tmp;
  auto up = [&](const void *h, size_t nb) -> void * {
    if (!h) return nullptr;
    void *d = nullptr;
    if (hipMalloc(&d, nb) != hipSuccess) return nullptr;
    tmp.push_back(d);
    if (hipMemcpyAsync(d, h, nb, hipMemcpyHostToDevice, c->stream) != hipSuccess) return nullptr;
    return d;
  };
  std::vector<int64_t> tix(mp);
  for (size_t j = 0; j < mp; j++) tix[j] = c->order.empty() ? (int64_t)j : c->order[j];
  MetArgs a;
  memset(&a, 0, sizeof a);
  const size_t fb = sizeof(float) * n;
  a.swdown = (const float *)up(m->swdown, fb);
  a.lwdown = (const float *)up(m->lwdown, fb);
  a.tair = (const float *)up(m->tair, fb);
  a.qair = (const float *)up(m->qair, fb);
  a.wind = (const float *)up(m->wind, fb);
  a.wind_e = (const float *)up(m->wind_e, fb);
  a.psurf = (const float *)up(m->psurf, fb);
  a.rainf = (const float *)up(m->rainf, fb);
  a.tix = (const int64_t *)up(tix.data(), sizeof(int64_t) * mp);
  a.elev = (const double *)up(m->elevation, sizeof(double) * mp);
  a.lat = (const double *)up(m->latitude, sizeof(double) * mp);
  a.hod = (const double *)up(hod.data(), sizeof(double) * hod.size());
  a.doy = (const double *)up(doy.data(), sizeof(double) * doy.size());
  a.tair_offset = m->tair_offset;
  a.psurf_scale = m->psurf_scale;
  a.rainf_scale = m->rainf_scale;
  a.qair_is_rh = m->qair_is_rh;
  const bool ok = a.swdown && a.tair && a.qair && a.wind && a.rainf && a.tix && a.lat && a.hod && a.doy &&
                  (!m->lwdown || a.lwdown) && (!m->wind_e || a.wind_e) && (!m->psurf || a.psurf) &&
                  (!m->elevation || a.elev);
  hipError_t e = hipSuccess;
  if (ok) {
    const dim3 grid((unsigned)((mp + 255) / 256), (unsigned)nsteps);
    void **f = c->forc.data();
    if (c->precision == CABLE_PREC_MIXED)
      hipLaunchKernelGGL(met_kernel<float>, grid, dim3(256), 0, c->stream, c->mp, a, (float *)f[0], (float *)f[1],
                         (float *)f[2], (float *)f[3], (float *)f[4], (float *)f[5], (float *)f[6], (float *)f[7]);
    else
      hipLaunchKernelGGL(met_kernel<double>, grid, dim3(256), 0, c->stream, c->mp, a, (double *)f[0], (double *)f[1],
                         (double *)f[2], (double *)f[3], (double *)f[4], (double *)f[5], (double *)f[6],
                         (double *)f[7]);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  for (void *d : tmp) hipFree(d);
  if (!ok) return fail(CABLE_E_HIP, "cable_gpu_set_met: device buffers for the met variables");
  if (e != hipSuccess) return fail(CABLE_E_HIP, "cable_gpu_set_met: %s", hipGetErrorString(e));
  const int ids[8] = {CABLE_F_met_fsd, CABLE_F_met_fld, CABLE_F_met_tk, CABLE_F_met_qv,
                      CABLE_F_met_ua, CABLE_F_met_pmb, CABLE_F_met_precip, CABLE_F_met_coszen};
  std::vector<void *> tab = c->dev;
  for (int i = 0; i < 8; i++) tab[ids[i]] = c->forc[i];
  HIPCHECK(hipMemcpyAsync(c->d_tab_run, tab.data(), sizeof(void *) * CABLE_NFIELDS, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->nsteps_forc = nsteps;
  c->forc_k0 = k_first;
  c->forc_doy = doy;
  c->ca = m->fixed_co2 / 1000000.0;
  c->doy = doy[0];
  return CABLE_OK;
}

int cable_gpu_met_clock(int precision, double shod, int sdoy, int smoy, int syear, double dels, int k_first,
                        int nsteps, double *hod, double *doy) {
  if (!hod || !doy || k_first < 0 || nsteps <= 0 || !(dels > 0)) return fail(CABLE_E_ARG, "cable_gpu_met_clock: bad arguments");
  std::vector<double> h, d;
  if (precision == CABLE_PREC_MIXED)
    met_clock<float>(shod, sdoy, smoy, syear, dels, k_first, nsteps, h, d);
  else
    met_clock<double>(shod, sdoy, smoy, syear, dels, k_first, nsteps, h, d);
  memcpy(hod, h.data(), sizeof(double) * (size_t)nsteps);
  memcpy(doy, d.data(), sizeof(double) * (size_t)nsteps);
  return CABLE_OK;
}

int cable_gpu_download_forcing(cable_gpu_ctx *c, int k, int which, void *host, int host_type) {
  if (!c || !host || which < 0 || which >= 8) return fail(CABLE_E_ARG, "cable_gpu_download_forcing: bad arguments");
  if (c->nsteps_forc == 0 || k < c->forc_k0 || k >= c->forc_k0 + c->nsteps_forc)
    return fail(CABLE_E_ARG, "cable_gpu_download_forcing: step %d is not resident", k);
  const size_t mp = (size_t)c->mp;
  FieldMeta fm = {"forcing", 'F', 1, ROLE_FORC};
  const size_t de = dev_elem(fm, c->precision);
  std::vector<unsigned char> buf(de * mp);
  HIPCHECK(hipMemcpyAsync(buf.data(), (unsigned char *)c->forc[which] + de * mp * (size_t)(k - c->forc_k0),
                          buf.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  const int64_t *ord = c->order.empty() ? nullptr : c->order.data();
  for (size_t j = 0; j < mp; j++) {
    const size_t d = ord ? (size_t)ord[j] : j;
    const double v = de == 8 ? ((double *)buf.data())[j] : (double)((float *)buf.data())[j];
    if (host_type == CABLE_HOST_F64)
      ((double *)host)[d] = v;
    else if (host_type == CABLE_HOST_F32)
      ((float *)host)[d] = (float)v;
    else
      return fail(CABLE_E_ARG, "cable_gpu_download_forcing: bad host type %d", host_type);
  }
  return CABLE_OK;
}

int cable_gpu_consistency_check(cable_gpu_ctx *c, int enable) {
  if (!c) return fail(CABLE_E_ARG, "cable_gpu_consistency_check: null context");
  if (enable) {
    if (!c->cons_part) HIPCHECK(hipMalloc(&c->cons_part, sizeof(double) * 4 * (size_t)step_workgroups(c)));
    if (!c->cons_acc) HIPCHECK(hipMalloc(&c->cons_acc, sizeof(double) * 4));
    HIPCHECK(hipMemsetAsync(c->cons_acc, 0, sizeof(double) * 4, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
  }
  c->cons_on = enable != 0;
  return CABLE_OK;
}

int cable_gpu_consistency_values(cable_gpu_ctx *c, double *out) {
  if (!c || !out) return fail(CABLE_E_ARG, "cable_gpu_consistency_values: null argument");
  if (!c->cons_acc) return fail(CABLE_E_STATE, "cable_gpu_consistency_values: the check was never enabled");
  double *tot = nullptr;
  HIPCHECK(hipMalloc(&tot, sizeof(double) * 2));
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(CONS_THREADS), 0, c->stream, (const double *)c->dev[CABLE_F_bal_ebal_tot],
                     c->mp, tot);
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(CONS_THREADS), 0, c->stream, (const double *)c->dev[CABLE_F_bal_wbal_tot],
                     c->mp, tot + 1);
  HIPCHECK(hipGetLastError());
  double acc[4], t2[2];
  HIPCHECK(hipMemcpyAsync(acc, c->cons_acc, sizeof acc, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipMemcpyAsync(t2, tot, sizeof t2, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipFree(tot));
  const double n = acc[0], mp = (double)c->mp;
  out[0] = n;
  out[1] = acc[1];
  out[2] = n > 0 ? acc[3] / n : 0.0;
  out[3] = n > 0 ? acc[2] / n : 0.0;
  out[4] = n > 0 ? t2[0] / mp / n : 0.0;
  out[5] = n > 0 ? t2[1] / mp / n : 0.0;
  return CABLE_OK;
}

int cable_gpu_run(cable_gpu_ctx *c, int k0, int nsteps, double dels) {
  if (!c) return fail(CABLE_E_ARG, "cable_gpu_run: null context");
  if (!c->have_tables || c->nsteps_forc == 0)
    return fail(CABLE_E_STATE, "cable_gpu_run: class tables and forcing must be set first");
  if (k0 < c->forc_k0 || nsteps <= 0 || k0 + nsteps > c->forc_k0 + c->nsteps_forc || !(dels > 0))
    return fail(CABLE_E_ARG, "cable_gpu_run: steps %d..%d outside the resident forcing steps %d..%d", k0,
                k0 + nsteps - 1, c->forc_k0, c->forc_k0 + c->nsteps_forc - 1);
  HIPCHECK(hipMemsetAsync(c->d_npass, 0, sizeof(unsigned long long), c->stream));
  c->hyb_valid = false;   // the plain step does not keep the surrogate's inputs
  HIPCHECK(hipEventRecord(c->ev0, c->stream));
  for (int k = k0; k < k0 + nsteps; k++) {
    StepArgs a = make_args(c, c->d_tab_run, k + 1, dels, (int64_t)(k - c->forc_k0) * c->mp);
    if (c->cons_on) a.cons_part = c->cons_part;
    hipError_t e = launch(c, a, false);
    if (e == hipSuccess && c->cons_on) e = launch_cons_reduce(c);
    if (e != hipSuccess) return fail(CABLE_E_HIP, "cable_gpu_run launch: %s", hipGetErrorString(e));
    c->first_call = false;
  }
  HIPCHECK(hipEventRecord(c->ev1, c->stream));
  HIPCHECK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  c->last_step_ms = ms;
  c->last_infer_ms = 0.0;
  unsigned long long np = 0;
  HIPCHECK(hipMemcpy(&np, c->d_npass, sizeof np, hipMemcpyDeviceToHost));
  c->last_passes = (int64_t)np;
  return CABLE_OK;
}

// ---------------------------------------------------------------------------
// surrogate: capture, weights, hybrid run

namespace {
int check_resident(cable_gpu_ctx *c, int k0, int nsteps, double dels, const char *who) {
  if (!c) return fail(CABLE_E_ARG, "%s: null context", who);
  if (c->precision != CABLE_PREC_MIXED) return fail(CABLE_E_STATE, "%s: needs a mixed-precision context", who);
  if (!c->have_tables || c->nsteps_forc == 0)
    return fail(CABLE_E_STATE, "%s: class tables and forcing must be set first", who);
  if (k0 < c->forc_k0 || nsteps <= 0 || k0 + nsteps > c->forc_k0 + c->nsteps_forc || !(dels > 0))
    return fail(CABLE_E_ARG, "%s: steps %d..%d outside the resident forcing steps %d..%d", who, k0,
                k0 + nsteps - 1, c->forc_k0, c->forc_k0 + c->nsteps_forc - 1);
  return CABLE_OK;
}

int ensure_hyb(cable_gpu_ctx *c) {
  if (!c->hyb) {
    HIPCHECK(hipMalloc(&c->hyb, sizeof(float) * 3 * (size_t)c->mp));
    c->hyb_valid = false;
  }
  return CABLE_OK;
}
}  // namespace

int cable_gpu_run_capture(cable_gpu_ctx *c, int k0, int nsteps, double dels, int64_t capacity) {
  int r = check_resident(c, k0, nsteps, dels, "cable_gpu_run_capture");
  if (r) return r;
  if (capacity <= 0) return fail(CABLE_E_ARG, "cable_gpu_run_capture: capacity %lld", (long long)capacity);
  if ((r = ensure_hyb(c))) return r;
  if (capacity > c->scap) {
    if (c->sx) HIPCHECK(hipFree(c->sx));
    if (c->sy) HIPCHECK(hipFree(c->sy));
    c->sx = c->sy = nullptr;
    c->scap = 0;
    HIPCHECK(hipMalloc(&c->sx, sizeof(float) * cbm_mlp::NIN * (size_t)capacity));
    HIPCHECK(hipMalloc(&c->sy, sizeof(float) * cbm_mlp::NOUT * (size_t)capacity));
    c->scap = capacity;
  }
  if (!c->scount) HIPCHECK(hipMalloc(&c->scount, sizeof(unsigned long long)));
  HIPCHECK(hipMemsetAsync(c->scount, 0, sizeof(unsigned long long), c->stream));
  HIPCHECK(hipMemsetAsync(c->d_npass, 0, sizeof(unsigned long long), c->stream));
  for (int k = k0; k < k0 + nsteps; k++) {
    StepArgs a = make_args(c, c->d_tab_run, k + 1, dels, (int64_t)(k - c->forc_k0) * c->mp);
    a.hyb = c->hyb;
    a.hyb_valid = c->hyb_valid ? 1 : 0;
    a.sx = c->sx;
    a.sy = c->sy;
    a.scount = c->scount;
    a.scap = (unsigned long long)c->scap;
    hipError_t e = launch_mixed_mode(MODE_CAPTURE, a, c->mp, c->stream);
    if (e != hipSuccess) return fail(CABLE_E_HIP, "cable_gpu_run_capture launch: %s", hipGetErrorString(e));
    c->first_call = false;
    c->hyb_valid = true;
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  unsigned long long np = 0;
  HIPCHECK(hipMemcpy(&np, c->d_npass, sizeof np, hipMemcpyDeviceToHost));
  c->last_passes = (int64_t)np;
  return CABLE_OK;
}

int cable_gpu_samples(cable_gpu_ctx *c, const float **x, const float **y, int64_t *n) {
  if (!c || !x || !y || !n) return fail(CABLE_E_ARG, "cable_gpu_samples: null argument");
  if (!c->scount) return fail(CABLE_E_STATE, "cable_gpu_samples: nothing captured");
  unsigned long long k = 0;
  HIPCHECK(hipMemcpy(&k, c->scount, sizeof k, hipMemcpyDeviceToHost));
  *x = c->sx;
  *y = c->sy;
  *n = (int64_t)(k < (unsigned long long)c->scap ? k : (unsigned long long)c->scap);
  return CABLE_OK;
}

int cable_gpu_copy_samples(cable_gpu_ctx *c, int64_t first, int64_t n, float *x, float *y) {
  if (!c || !x || !y || first < 0 || n <= 0) return fail(CABLE_E_ARG, "cable_gpu_copy_samples: bad arguments");
  if (first + n > c->scap) return fail(CABLE_E_ARG, "cable_gpu_copy_samples: rows beyond the capture buffer");
  HIPCHECK(hipMemcpy(x, c->sx + first * cbm_mlp::NIN, sizeof(float) * cbm_mlp::NIN * (size_t)n, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(y, c->sy + first * cbm_mlp::NOUT, sizeof(float) * cbm_mlp::NOUT * (size_t)n, hipMemcpyDeviceToHost));
  return CABLE_OK;
}

int cable_gpu_set_surrogate(cable_gpu_ctx *c, const float *params, const float *norm) {
  using namespace cbm_mlp;
  if (!c || !params || !norm) return fail(CABLE_E_ARG, "cable_gpu_set_surrogate: null argument");
  for (int k = 0; k < NIN; k++)
    if (!(norm[NIN + k] > 0)) return fail(CABLE_E_ARG, "cable_gpu_set_surrogate: sd_x[%d] not positive", k);
  // weights transposed to [out][in] and rounded to BF16 for the matrix cores
  std::vector<uint16_t> wt;
  for (int l = 0; l < 4; l++)
    for (int j = 0; j < LAYER_OUT[l]; j++)
      for (int i = 0; i < LAYER_IN[l]; i++) wt.push_back(f32_to_bf16(params[LAYER_W[l] + i * LAYER_OUT[l] + j]));
  if (!c->sparams) HIPCHECK(hipMalloc(&c->sparams, sizeof(float) * NPARAM));
  if (!c->snorm) HIPCHECK(hipMalloc(&c->snorm, sizeof(float) * NNORM));
  if (!c->swt) HIPCHECK(hipMalloc(&c->swt, sizeof(uint16_t) * wt.size()));
  HIPCHECK(hipMemcpy(c->sparams, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->snorm, norm, sizeof(float) * NNORM, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->swt, wt.data(), sizeof(uint16_t) * wt.size(), hipMemcpyHostToDevice));
  // the previous step's leaf state the network reads (also kept by the
  // allocation-free hybrid run)
  int r = ensure_hyb(c);
  if (r) return r;
  return CABLE_OK;
}

int cable_gpu_run_hybrid(cable_gpu_ctx *c, int k0, int nsteps, double dels, int max_passes, double resid_max,
                         int use_surrogate) {
  int r = check_resident(c, k0, nsteps, dels, "cable_gpu_run_hybrid");
  if (r) return r;
  if (use_surrogate && !c->swt) return fail(CABLE_E_STATE, "cable_gpu_run_hybrid: no surrogate set");
  if (max_passes < 1 || max_passes > 20) return fail(CABLE_E_ARG, "cable_gpu_run_hybrid: max_passes %d", max_passes);
  if ((r = ensure_hyb(c))) return r;
  if (!c->pred) {
    // zero-filled: the inference kernel skips workgroups without leaves
    HIPCHECK(hipMalloc(&c->pred, sizeof(float) * cbm_mlp::NOUT * (size_t)c->mp));
    HIPCHECK(hipMemset(c->pred, 0, sizeof(float) * cbm_mlp::NOUT * (size_t)c->mp));
  }
  if (!c->nfallback) HIPCHECK(hipMalloc(&c->nfallback, sizeof(unsigned long long)));
  if (!c->d_lit) HIPCHECK(hipMalloc(&c->d_lit, sizeof(unsigned long long)));
  const int64_t nlitblk = (c->mp + 255) / 256;
  if (!c->d_litblk) HIPCHECK(hipMalloc(&c->d_litblk, sizeof(unsigned long long) * (size_t)nlitblk));
  while (c->kev.size() < 4 * (size_t)nsteps) {
    hipEvent_t ev;
    HIPCHECK(hipEventCreate(&ev));
    c->kev.push_back(ev);
  }
  HIPCHECK(hipMemsetAsync(c->nfallback, 0, sizeof(unsigned long long), c->stream));
  HIPCHECK(hipMemsetAsync(c->d_npass, 0, sizeof(unsigned long long), c->stream));
  HIPCHECK(hipEventRecord(c->ev0, c->stream));
  unsigned long long lit = 0;
  for (int k = k0; k < k0 + nsteps; k++) {
    StepArgs a = make_args(c, c->d_tab_run, k + 1, dels, (int64_t)(k - c->forc_k0) * c->mp);
    if ((k - k0) % LIT_CHUNK == 0) {
      // the steps of the next chunk with a leaf-bearing tile in daylight (one
      // scan of their shortwave and one wait per 64 steps)
      const int nst = std::min(LIT_CHUNK, k0 + nsteps - k);
      hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, c->stream, c->d_lit, 1);
      HIPCHECK(hipMemsetAsync(c->d_litblk, 0, sizeof(unsigned long long) * (size_t)nlitblk, c->stream));
      const int64_t blocks = std::min<int64_t>((c->mp + 255) / 256, 2048);
      hipLaunchKernelGGL(lit_steps_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, a, nst, c->d_lit,
                         c->d_litblk);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpyAsync(&lit, c->d_lit, sizeof lit, hipMemcpyDeviceToHost, c->stream));
      HIPCHECK(hipStreamSynchronize(c->stream));
    }
    // a step with no leaf-bearing tile in daylight warm-starts nothing: it
    // runs no network and the plain step (the same bits as the hybrid step
    // there).  Before a daylight step, or where the next step is not known
    // (the chunk's or the run's last), that step also keeps the leaf state
    // the next step's features read.
    const int s = (k - k0) % LIT_CHUNK;
    const bool day = (lit >> s) & 1ull;
    const bool next_day = s + 1 == LIT_CHUNK || k + 1 == k0 + nsteps || ((lit >> (s + 1)) & 1ull);
    const int mode = day ? MODE_HYBRID : next_day ? MODE_UNLIT : MODE_PLAIN;
    a.hyb = c->hyb;
    a.hyb_valid = c->hyb_valid ? 1 : 0;
    a.norm = c->snorm;
    a.pred = use_surrogate ? c->pred : nullptr;
    a.max_passes = max_passes;
    a.resid_max = (float)resid_max;
    a.nfallback = c->nfallback;
    a.lit_blocks = c->d_litblk;
    a.lit_bit = s;
    hipError_t e = hipSuccess;
    hipEvent_t *ev = &c->kev[4 * (size_t)(k - k0)];
    if (c->launch_timing) HIPCHECK(hipEventRecord(ev[0], c->stream));
    if (use_surrogate && day) e = launch_mlp_infer(a, c->swt, c->sparams, c->pred, nullptr, c->mp, c->stream);
    if (c->launch_timing) HIPCHECK(hipEventRecord(ev[1], c->stream));
    // the capped step; a tile the residual test rejects falls back to the
    // reference's iteration within the same launch
    if (c->launch_timing) HIPCHECK(hipEventRecord(ev[2], c->stream));
    if (e == hipSuccess) e = launch_mixed_mode(mode, a, c->mp, c->stream);
    if (c->launch_timing) HIPCHECK(hipEventRecord(ev[3], c->stream));
    if (e != hipSuccess) return fail(CABLE_E_HIP, "cable_gpu_run_hybrid launch: %s", hipGetErrorString(e));
    c->first_call = false;
    c->hyb_valid = true;
  }
  HIPCHECK(hipEventRecord(c->ev1, c->stream));
  HIPCHECK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  unsigned long long np = 0, fb = 0;
  HIPCHECK(hipMemcpy(&np, c->d_npass, sizeof np, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(&fb, c->nfallback, sizeof fb, hipMemcpyDeviceToHost));
  c->last_infer_ms = c->last_step_ms = 0.0;
  if (!c->launch_timing) c->last_step_ms = ms;
  for (int k = 0; c->launch_timing && k < nsteps; k++) {
    float a = 0.f, b = 0.f;
    HIPCHECK(hipEventElapsedTime(&a, c->kev[4 * (size_t)k], c->kev[4 * (size_t)k + 1]));
    HIPCHECK(hipEventElapsedTime(&b, c->kev[4 * (size_t)k + 2], c->kev[4 * (size_t)k + 3]));
    c->last_infer_ms += a;
    c->last_step_ms += b;
  }
  c->last_passes = (int64_t)np;
  c->last_fallbacks = (int64_t)fb;
  c->last_tile_steps = (int64_t)nsteps * c->mp;
  return CABLE_OK;
}

int cable_gpu_run_hybrid_noalloc(cable_gpu_ctx *c, int k0, int nsteps, double dels, int max_passes,
                                 double resid_max, int use_surrogate) {
  int r = check_resident(c, k0, nsteps, dels, "cable_gpu_run_hybrid_noalloc");
  if (r) return r;
  if (c->mp > NOALLOC_MAX_TILES)
    return fail(CABLE_E_ARG, "cable_gpu_run_hybrid_noalloc: %lld tiles, at most %lld (the size of its scratch)",
                (long long)c->mp, (long long)NOALLOC_MAX_TILES);
  if (use_surrogate && !c->swt) return fail(CABLE_E_STATE, "cable_gpu_run_hybrid_noalloc: no surrogate set");
  if (!c->hyb)
    return fail(CABLE_E_STATE, "cable_gpu_run_hybrid_noalloc: no surrogate state (call cable_gpu_set_surrogate)");
  if (max_passes < 1 || max_passes > 20)
    return fail(CABLE_E_ARG, "cable_gpu_run_hybrid_noalloc: max_passes %d", max_passes);
  void *base = g_scratch[c->device];
  if (!base) return fail(CABLE_E_STATE, "cable_gpu_run_hybrid_noalloc: no scratch on device %d", c->device);
  float *pred = (float *)base;
  unsigned long long *cnt = (unsigned long long *)((char *)base + NOALLOC_PRED_BYTES);  // passes, fallbacks
  hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, c->stream, cnt, 2);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipEventRecord(c->ev0, c->stream));
  for (int k = k0; k < k0 + nsteps; k++) {
    StepArgs a = make_args(c, c->d_tab_run, k + 1, dels, (int64_t)(k - c->forc_k0) * c->mp);
    a.npass = cnt;
    a.hyb = c->hyb;
    a.hyb_valid = c->hyb_valid ? 1 : 0;
    a.norm = c->snorm;
    a.pred = use_surrogate ? pred : nullptr;
    a.max_passes = max_passes;
    a.resid_max = (float)resid_max;
    a.nfallback = cnt + 1;
    hipError_t e = hipSuccess;
    if (use_surrogate) e = launch_mlp_infer(a, c->swt, c->sparams, pred, nullptr, c->mp, c->stream);
    if (e == hipSuccess) e = launch_mixed_mode(MODE_HYBRID, a, c->mp, c->stream);
    if (e != hipSuccess) return fail(CABLE_E_HIP, "cable_gpu_run_hybrid_noalloc launch: %s", hipGetErrorString(e));
    c->first_call = false;
    c->hyb_valid = true;
  }
  HIPCHECK(hipEventRecord(c->ev1, c->stream));
  HIPCHECK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  unsigned long long h[2] = {0ull, 0ull};
  HIPCHECK(hipMemcpy(h, cnt, sizeof h, hipMemcpyDeviceToHost));
  c->last_ms = ms;
  c->last_step_ms = ms;
  c->last_infer_ms = 0.0;
  c->last_passes = (int64_t)h[0];
  c->last_fallbacks = (int64_t)h[1];
  c->last_tile_steps = (int64_t)nsteps * c->mp;
  return CABLE_OK;
}

int cable_gpu_surrogate_eval(cable_gpu_ctx *c, int k, float *x, float *pred) {
  int r = check_resident(c, k, 1, 1.0, "cable_gpu_surrogate_eval");
  if (r) return r;
  if (!c->swt) return fail(CABLE_E_STATE, "cable_gpu_surrogate_eval: no surrogate set");
  if (!x || !pred) return fail(CABLE_E_ARG, "cable_gpu_surrogate_eval: null output");
  using namespace cbm_mlp;
  const size_t mp = (size_t)c->mp;
  float *dx = nullptr, *dp = nullptr;
  HIPCHECK(hipMalloc(&dx, sizeof(float) * NIN * mp));
  HIPCHECK(hipMalloc(&dp, sizeof(float) * NOUT * mp));
  StepArgs a = make_args(c, c->d_tab_run, k + 1, 1.0, (int64_t)(k - c->forc_k0) * c->mp);
  a.hyb = c->hyb;
  a.hyb_valid = (c->hyb && c->hyb_valid) ? 1 : 0;
  a.norm = c->snorm;
  hipError_t e = launch_mlp_infer(a, c->swt, c->sparams, dp, dx, c->mp, c->stream);
  if (e != hipSuccess) return fail(CABLE_E_HIP, "cable_gpu_surrogate_eval launch: %s", hipGetErrorString(e));
  std::vector<float> hx(NIN * mp), hp(NOUT * mp);
  HIPCHECK(hipMemcpyAsync(hx.data(), dx, sizeof(float) * hx.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipMemcpyAsync(hp.data(), dp, sizeof(float) * hp.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipFree(dx));
  HIPCHECK(hipFree(dp));
  // rows back to the caller's tile order
  for (size_t j = 0; j < mp; j++) {
    const size_t d = c->order.empty() ? j : (size_t)c->order[j];
    memcpy(x + d * NIN, hx.data() + j * NIN, sizeof(float) * NIN);
    memcpy(pred + d * NOUT, hp.data() + j * NOUT, sizeof(float) * NOUT);
  }
  return CABLE_OK;
}

int cable_gpu_hybrid_counters(cable_gpu_ctx *c, int64_t *fallback_tile_steps, int64_t *tile_steps) {
  if (!c || !fallback_tile_steps || !tile_steps) return fail(CABLE_E_ARG, "cable_gpu_hybrid_counters: null argument");
  *fallback_tile_steps = c->last_fallbacks;
  *tile_steps = c->last_tile_steps;
  return CABLE_OK;
}

int cable_gpu_sync(cable_gpu_ctx *c) {
  if (!c) return fail(CABLE_E_ARG, "cable_gpu_sync: null context");
  HIPCHECK(hipStreamSynchronize(c->stream));
  return CABLE_OK;
}

int cable_gpu_last_run_ms(cable_gpu_ctx *c, double *ms) {
  if (!c || !ms) return fail(CABLE_E_ARG, "cable_gpu_last_run_ms: null argument");
  *ms = c->last_ms;
  return CABLE_OK;
}

int cable_gpu_set_launch_timing(cable_gpu_ctx *c, int on) {
  if (!c) return fail(CABLE_E_ARG, "cable_gpu_set_launch_timing: null context");
  c->launch_timing = on != 0;
  return CABLE_OK;
}

int cable_gpu_last_kernel_ms(cable_gpu_ctx *c, double *step_ms, double *infer_ms) {
  if (!c || !step_ms || !infer_ms) return fail(CABLE_E_ARG, "cable_gpu_last_kernel_ms: null argument");
  *step_ms = c->last_step_ms;
  *infer_ms = c->last_infer_ms;
  return CABLE_OK;
}

int cable_gpu_counters(cable_gpu_ctx *c, int64_t *passes) {
  if (!c || !passes) return fail(CABLE_E_ARG, "cable_gpu_counters: null argument");
  *passes = c->last_passes;
  return CABLE_OK;
}

}  // extern "C"

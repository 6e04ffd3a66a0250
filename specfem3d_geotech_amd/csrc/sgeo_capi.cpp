// sgeo_capi.cpp -- host orchestration and C ABI (include/sgeo.h) of the
// MI355X EBE-PCG + viscoplastic hot path.  One context = one HIP device +
// one stream; all device work of a context is ordered on that stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sgeo.h"
#include "sgeo_internal.h"

using namespace sgeo;

// --------------------------------------------------------------------------
// errors
// --------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(call)                                                        \
  do {                                                                      \
    hipError_t _e = (call);                                                 \
    if (_e != hipSuccess)                                                   \
      return fail(SGEO_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,     \
                  hipGetErrorString(_e));                                   \
  } while (0)

extern "C" const char* sgeo_last_error(void) { return g_err.c_str(); }

// --------------------------------------------------------------------------
// objects
// --------------------------------------------------------------------------
struct sgeo_ctx {
  int device = 0;
  int n_cu = 0;  // compute units of the device
  hipStream_t stream = nullptr;
  PcgScalars* d_sc = nullptr;
  PcgScalars* h_sc = nullptr;  // pinned
  PlasticScalars* d_ps = nullptr;
  PlasticScalars* h_ps = nullptr;  // pinned
  double* d_part = nullptr;
  size_t part_cap = 0;  // doubles
  // Seam-1 cache
  sgeo_mesh* cache = nullptr;
  const void* cache_k = nullptr;
  const void* cache_g = nullptr;
  const void* cache_d = nullptr;
  uint64_t cache_sig = 0;
  double* d_u = nullptr;
  double* d_f = nullptr;
  size_t uf_cap = 0;
  int reduction = SGEO_REDUCE_TREE;
  // this device's copy of the allocation-free entry point's scratch array
  // (hipGetSymbolAddress at sgeo_init), and a pinned word for its checks
  unsigned char* d_scratch = nullptr;
  int32_t* h_flag = nullptr;
  // HIP-event timing of the element-product kernel launches in PCG loops
  bool prof = false;
  std::vector<hipEvent_t> ev;  // pairs, reused per chunk
  int ev_used = 0;
  int64_t prof_count = 0;
  double prof_ms = 0.0;
};

static int prof_flush(sgeo_ctx* c) {
  // events of completed launches (caller has synchronized the stream)
  for (int i = 0; i + 1 < c->ev_used; i += 2) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
    c->prof_ms += ms;
    c->prof_count++;
  }
  c->ev_used = 0;
  return 0;
}

static int prof_mark(sgeo_ctx* c) {
  if (!c->prof) return 0;
  if ((size_t)c->ev_used >= c->ev.size()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    c->ev.push_back(e);
  }
  HIPCHK(hipEventRecord(c->ev[c->ev_used++], c->stream));
  return 0;
}

struct sgeo_mesh {
  sgeo_ctx* ctx = nullptr;
  int32_t nelmt = 0, neq = 0, nmat = 0;
  int32_t* d_gdof = nullptr;
  int32_t* d_rowptr = nullptr;
  int32_t* d_slots = nullptr;
  double* d_corners = nullptr;
  int32_t* d_mat = nullptr;
  int kprec = -1;  // -1: no stiffness yet
  int kfmt = SGEO_KFMT_FULL;
  double* d_K64 = nullptr;  // canonical: FP64, reference layout
  float* d_K32 = nullptr;   // derived operators (see refresh_operator)
  double* d_KP64 = nullptr;
  float* d_KP32 = nullptr;
  int k_ver = 0, k32_ver = -1, kp64_ver = -1, kp32_ver = -1;
  double* d_dinv = nullptr;
  // node map of the nodal result fields (sgeo_mesh_set_nodes)
  int32_t nnode = 0;
  int32_t* d_nrowptr = nullptr;
  int32_t* d_nslots = nullptr;
  // work vectors
  double* d_Y = nullptr;
  double *d_r = nullptr, *d_z = nullptr, *d_p = nullptr, *d_kp = nullptr;
  size_t bytes = 0;
};

struct sgeo_slope {
  sgeo_mesh* mesh = nullptr;
  double *d_x = nullptr, *d_oldx = nullptr, *d_load = nullptr;
  double *d_body = nullptr, *d_ext = nullptr;
  double *d_evpt = nullptr, *d_stress = nullptr, *d_BL = nullptr;
  double* d_vm = nullptr;  // excavation: per-point von Mises plastic strain
  double* d_sf = nullptr;  // excavation: per-point stress concentration factor
  double* d_nacc = nullptr;  // excavation: nodal sums of vm, minima of sf (2 x nnode)
  int32_t nacc_nnode = 0;    // node count d_nacc was sized for
  int nacc_init = 0;         // d_nacc holds a started accumulation
  uint8_t* d_flags = nullptr;
  MatParams* d_mp = nullptr;
  std::vector<MatParams> h_mp;
  bool begun = false;
  int excav = 0;
};

template <typename T>
static int dalloc(T** p, size_t n, size_t* acc = nullptr) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess)
    return fail(SGEO_ENOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                hipGetErrorString(e));
  if (acc) *acc += n * sizeof(T);
  return 0;
}

#define TRY(x)               \
  do {                       \
    int _r = (x);            \
    if (_r != 0) return _r;  \
  } while (0)

// Device temporaries of one call, freed when the call returns, on every
// path (hipFree waits for the device).
struct DevScratch {
  std::vector<void*> ptrs;
  ~DevScratch() {
    for (void* p : ptrs)
      if (p) hipFree(p);
  }
  template <typename T>
  int alloc(T** p, size_t n) {
    const int r = dalloc(p, n);
    if (r == 0) ptrs.push_back((void*)*p);
    return r;
  }
};

// Destroys a handle under construction when a create call returns early
// with an error (every destroy function frees only what was allocated);
// release() hands it to the caller on success.
template <typename T, int (*Destroy)(T*)>
struct Partial {
  T* p;
  ~Partial() {
    if (p) Destroy(p);
  }
  T* release() {
    T* q = p;
    p = nullptr;
    return q;
  }
};

static int ensure_part(sgeo_ctx* c, size_t n) {
  if (n <= c->part_cap) return 0;
  if (c->d_part) hipFree(c->d_part);
  c->d_part = nullptr;
  c->part_cap = 0;
  TRY(dalloc(&c->d_part, n));
  c->part_cap = n;
  return 0;
}

// --------------------------------------------------------------------------
// scratch of the allocation-free drop-in (sgeo_pcg_solver_noalloc): device
// memory of a fixed size that the code object brings with it, one copy per
// GPU, held until the process ends.  Its name is a host-side stand-in; host
// code uses only the per-device address hipGetSymbolAddress returns.
// --------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[SGEO_NOALLOC_SCRATCH_BYTES];

// --------------------------------------------------------------------------
// context
// --------------------------------------------------------------------------
extern "C" int sgeo_init(int device, sgeo_ctx** out) {
  if (!out) return fail(SGEO_EARG, "sgeo_init: null output");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fail(SGEO_EARG, "sgeo_init: device %d of %d", device, ndev);
  HIPCHK(hipSetDevice(device));
  sgeo_ctx* c = new sgeo_ctx();
  Partial<sgeo_ctx, sgeo_finalize> guard{c};
  c->device = device;
  HIPCHK(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device));
  if (c->n_cu < 1) c->n_cu = 1;
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  TRY(dalloc(&c->d_sc, 1));
  TRY(dalloc(&c->d_ps, 1));
  HIPCHK(hipHostMalloc((void**)&c->h_sc, sizeof(PcgScalars)));
  HIPCHK(hipHostMalloc((void**)&c->h_ps, sizeof(PlasticScalars)));
  HIPCHK(hipHostMalloc((void**)&c->h_flag, sizeof(int32_t)));
  // the device-current copy of the scratch array (every GPU has its own)
  HIPCHK(hipGetSymbolAddress((void**)&c->d_scratch, HIP_SYMBOL(scratch)));
  if (!c->d_scratch) return fail(SGEO_EHIP, "sgeo_init: no address for the scratch array");
  HIPCHK(hipMemsetAsync(c->d_sc, 0, sizeof(PcgScalars), c->stream));
  HIPCHK(hipMemsetAsync(c->d_ps, 0, sizeof(PlasticScalars), c->stream));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  *out = guard.release();
  return 0;
}

extern "C" int sgeo_finalize(sgeo_ctx* c) {
  if (!c) return 0;
  hipSetDevice(c->device);
  if (c->cache) sgeo_mesh_destroy(c->cache);
  if (c->d_u) hipFree(c->d_u);
  if (c->d_f) hipFree(c->d_f);
  if (c->d_part) hipFree(c->d_part);
  if (c->d_sc) hipFree(c->d_sc);
  if (c->d_ps) hipFree(c->d_ps);
  if (c->h_sc) hipHostFree(c->h_sc);
  if (c->h_ps) hipHostFree(c->h_ps);
  if (c->h_flag) hipHostFree(c->h_flag);
  for (hipEvent_t e : c->ev) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

extern "C" int sgeo_device_info(sgeo_ctx* c, char* name, size_t name_len,
                                int* n_cu, size_t* hbm) {
  if (!c) return fail(SGEO_EARG, "null ctx");
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, c->device));
  if (name && name_len) {
    snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (n_cu) *n_cu = p.multiProcessorCount;
  if (hbm) *hbm = p.totalGlobalMem;
  return 0;
}

extern "C" int sgeo_set_reduction(sgeo_ctx* c, int mode) {
  if (!c || (mode != SGEO_REDUCE_TREE && mode != SGEO_REDUCE_SEQUENTIAL))
    return fail(SGEO_EARG, "sgeo_set_reduction: mode %d", mode);
  c->reduction = mode;
  return 0;
}

extern "C" int sgeo_profile(sgeo_ctx* c, int enable, int64_t* launches,
                            double* total_ms) {
  if (!c) return fail(SGEO_EARG, "null ctx");
  if (launches) *launches = c->prof_count;
  if (total_ms) *total_ms = c->prof_ms;
  if (enable >= 0) {
    c->prof = enable != 0;
    c->prof_count = 0;
    c->prof_ms = 0.0;
    c->ev_used = 0;
  }
  return 0;
}

extern "C" int sgeo_synchronize(sgeo_ctx* c) {
  if (!c) return fail(SGEO_EARG, "null ctx");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// --------------------------------------------------------------------------
// mesh
// --------------------------------------------------------------------------
extern "C" int sgeo_mesh_create(sgeo_ctx* c, int32_t nelmt, int32_t neq,
                                const double* corners, const int32_t* gdof,
                                const int32_t* mat_id, int32_t nmat,
                                sgeo_mesh** out) {
  if (!c || !out || !gdof || nelmt <= 0 || neq <= 0)
    return fail(SGEO_EARG, "sgeo_mesh_create: bad arguments");
  if ((int64_t)nelmt * 81 > INT32_MAX)
    return fail(SGEO_EARG, "sgeo_mesh_create: nelmt*81 exceeds int32");
  const int64_t ns = (int64_t)nelmt * 81;
  // validate equations and build the inverse map (counting sort, stable in
  // element order)
  std::vector<int32_t> rowptr((size_t)neq + 2, 0);
  for (int64_t q = 0; q < ns; q++) {
    const int32_t g = gdof[q];
    if (g < 0 || g > neq)
      return fail(SGEO_EARG, "gdof_elmt[%lld]=%d outside 0..%d", (long long)q, g,
                  neq);
    rowptr[(size_t)g + 1]++;
  }
  for (size_t g = 1; g < rowptr.size(); g++) rowptr[g] += rowptr[g - 1];
  std::vector<int32_t> slots((size_t)ns), fill(rowptr.begin(), rowptr.end() - 1);
  for (int64_t q = 0; q < ns; q++) slots[(size_t)fill[(size_t)gdof[q]]++] = (int32_t)q;
  if (mat_id) {
    for (int32_t e = 0; e < nelmt; e++)
      if (mat_id[e] < 1 || mat_id[e] > nmat)
        return fail(SGEO_EARG, "mat_id[%d]=%d outside 1..%d", e, mat_id[e], nmat);
  }
  HIPCHK(hipSetDevice(c->device));
  sgeo_mesh* m = new sgeo_mesh();
  Partial<sgeo_mesh, sgeo_mesh_destroy> guard{m};
  m->ctx = c;
  m->nelmt = nelmt;
  m->neq = neq;
  m->nmat = nmat;
  const int64_t n1 = (int64_t)neq + 1;
  TRY(dalloc(&m->d_gdof, (size_t)ns, &m->bytes));
  TRY(dalloc(&m->d_rowptr, (size_t)neq + 2, &m->bytes));
  TRY(dalloc(&m->d_slots, (size_t)ns, &m->bytes));
  TRY(dalloc(&m->d_Y, (size_t)ns, &m->bytes));
  TRY(dalloc(&m->d_r, (size_t)n1, &m->bytes));
  TRY(dalloc(&m->d_z, (size_t)n1, &m->bytes));
  TRY(dalloc(&m->d_p, (size_t)n1, &m->bytes));
  TRY(dalloc(&m->d_kp, (size_t)n1, &m->bytes));
  TRY(dalloc(&m->d_dinv, (size_t)n1, &m->bytes));
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(m->d_gdof, gdof, ns * 4, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(m->d_rowptr, rowptr.data(), rowptr.size() * 4,
                        hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(m->d_slots, slots.data(), ns * 4, hipMemcpyHostToDevice, s));
  if (corners) {
    TRY(dalloc(&m->d_corners, (size_t)nelmt * 24, &m->bytes));
    HIPCHK(hipMemcpyAsync(m->d_corners, corners, (size_t)nelmt * 24 * 8,
                          hipMemcpyHostToDevice, s));
  }
  if (mat_id) {
    TRY(dalloc(&m->d_mat, (size_t)nelmt, &m->bytes));
    HIPCHK(hipMemcpyAsync(m->d_mat, mat_id, (size_t)nelmt * 4,
                          hipMemcpyHostToDevice, s));
  }
  TRY(ensure_part(c, 4 * (size_t)vec_reduce_blocks(n1) + 64));
  HIPCHK(hipStreamSynchronize(s));  // host vectors die at return
  *out = guard.release();
  return 0;
}

extern "C" int sgeo_mesh_destroy(sgeo_mesh* m) {
  if (!m) return 0;
  hipSetDevice(m->ctx->device);
  hipStreamSynchronize(m->ctx->stream);
  void* ptrs[] = {m->d_gdof, m->d_rowptr, m->d_slots, m->d_corners, m->d_mat,
                  m->d_K64,  m->d_K32,    m->d_KP64,  m->d_KP32,
                  m->d_dinv,  m->d_Y,       m->d_r,
                  m->d_z,    m->d_p,      m->d_kp,    m->d_nrowptr, m->d_nslots};
  for (void* p : ptrs)
    if (p) hipFree(p);
  delete m;
  return 0;
}

// Node map of the nodal result fields: for every node, its (element, GLL
// point) slots e*27+i in ascending element order (counting sort, stable),
// so that device sums over a node's elements run in the reference's order
// (src/slope.F90:424-430).
extern "C" int sgeo_mesh_set_nodes(sgeo_mesh* m, int32_t nnode, const int32_t* g_num) {
  if (!m || !g_num || nnode <= 0)
    return fail(SGEO_EARG, "sgeo_mesh_set_nodes: bad arguments");
  const int64_t ns = (int64_t)m->nelmt * 27;
  std::vector<int32_t> rowptr((size_t)nnode + 1, 0);
  for (int64_t q = 0; q < ns; q++) {
    const int32_t n = g_num[q];
    if (n < 0 || n >= nnode)
      return fail(SGEO_EARG, "g_num[%lld]=%d outside 0..%d", (long long)q, n, nnode - 1);
    rowptr[(size_t)n + 1]++;
  }
  for (size_t n = 1; n < rowptr.size(); n++) rowptr[n] += rowptr[n - 1];
  std::vector<int32_t> slots((size_t)ns), next(rowptr.begin(), rowptr.end() - 1);
  for (int64_t q = 0; q < ns; q++) slots[(size_t)next[(size_t)g_num[q]]++] = (int32_t)q;
  HIPCHK(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  HIPCHK(hipStreamSynchronize(s));
  if (m->d_nrowptr) hipFree(m->d_nrowptr);
  if (m->d_nslots) hipFree(m->d_nslots);
  m->d_nrowptr = m->d_nslots = nullptr;
  m->nnode = 0;
  TRY(dalloc(&m->d_nrowptr, rowptr.size(), &m->bytes));
  TRY(dalloc(&m->d_nslots, (size_t)ns, &m->bytes));
  HIPCHK(hipMemcpyAsync(m->d_nrowptr, rowptr.data(), rowptr.size() * 4,
                        hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(m->d_nslots, slots.data(), (size_t)ns * 4, hipMemcpyHostToDevice,
                        s));
  HIPCHK(hipStreamSynchronize(s));  // host vectors die at return
  m->nnode = nnode;
  return 0;
}

// Nodal result fields of one trial's stress state (device, FP64 [e][27][6])
// and, for an excavation stage, its accumulated plastic strain and stress
// concentration factor (d_acc, 2 x nnode, NULL otherwise); FP32 outputs in
// component blocks, any of them may be NULL.  Without d_acc, vmeps is zero
// and scf is inftol: the slope driver updates neither (src/slope.F90:249,
// 366-376).
static int nodal_fields(sgeo_mesh* m, const double* d_stress, const double* d_acc,
                        float* sig, float* psig, float* nsig, float* scf, float* vmeps) {
  if (!m->d_nrowptr || !m->d_nslots || m->nnode <= 0)
    return fail(SGEO_ESTATE, "no node map: call sgeo_mesh_set_nodes first");
  const size_t nn = (size_t)m->nnode;
  HIPCHK(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  float* d_out = nullptr;
  TRY(dalloc(&d_out, nn * 12));
  float *ds = d_out, *dp = d_out + 6 * nn, *dn = d_out + 9 * nn, *dv = d_out + 10 * nn,
        *dc = d_out + 11 * nn;
  launch_nodal_fields(s, m->nnode, m->d_nrowptr, m->d_nslots, d_stress, sig ? ds : nullptr,
                      psig ? dp : nullptr, nsig ? dn : nullptr);
  if (d_acc)
    launch_nodal_acc_out(s, m->nnode, m->d_nrowptr, d_acc, vmeps ? dv : nullptr,
                         scf ? dc : nullptr);
  hipError_t e = hipGetLastError();
  struct Out {
    float* host;
    const float* dev;
    size_t n;
  } outs[] = {{sig, ds, 6 * nn}, {psig, dp, 3 * nn}, {nsig, dn, nn},
              {d_acc ? vmeps : nullptr, dv, nn}, {d_acc ? scf : nullptr, dc, nn}};
  for (const Out& o : outs)
    if (e == hipSuccess && o.host)
      e = hipMemcpyAsync(o.host, o.dev, o.n * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  hipFree(d_out);
  if (e != hipSuccess) return fail(SGEO_EHIP, "nodal fields: %s", hipGetErrorString(e));
  if (!d_acc) {
    if (vmeps) std::memset(vmeps, 0, nn * 4);
    if (scf)
      for (size_t n = 0; n < nn; n++) scf[n] = 1.0e32f;  // real(inftol)
  }
  return 0;
}

// The nodal fields of a host stress state (E,27,6) on this mesh: the
// in-situ stress plotted as the excavation's step 0 (src/excavation.F90:
// 303-323).
extern "C" int sgeo_mesh_nodal_fields(sgeo_mesh* m, const double* stress_elmt, float* sig,
                                      float* psig, float* nsig) {
  if (!m || !stress_elmt) return fail(SGEO_EARG, "sgeo_mesh_nodal_fields: null argument");
  const size_t ns = (size_t)m->nelmt * 27 * 6;
  HIPCHK(hipSetDevice(m->ctx->device));
  double* d_st = nullptr;
  TRY(dalloc(&d_st, ns));
  hipError_t e = hipMemcpyAsync(d_st, stress_elmt, ns * 8, hipMemcpyHostToDevice,
                                m->ctx->stream);
  int rc = e == hipSuccess
               ? nodal_fields(m, d_st, nullptr, sig, psig, nsig, nullptr, nullptr)
               : fail(SGEO_EHIP, "sgeo_mesh_nodal_fields: %s", hipGetErrorString(e));
  hipStreamSynchronize(m->ctx->stream);
  hipFree(d_st);
  return rc;
}

extern "C" int sgeo_mesh_bytes(sgeo_mesh* m, size_t* bytes) {
  if (!m || !bytes) return fail(SGEO_EARG, "null argument");
  *bytes = m->bytes;
  return 0;
}

static int set_k_storage(sgeo_mesh* m, int kprec) {
  if (kprec != SGEO_K_F64 && kprec != SGEO_K_F32)
    return fail(SGEO_EARG, "kprec %d", kprec);
  const size_t nk = (size_t)m->nelmt * 6561;
  if (!m->d_K64) TRY(dalloc(&m->d_K64, nk, &m->bytes));
  m->kprec = kprec;
  return 0;
}

// Derived storage of the selected operator, rebuilt from the canonical FP64
// full K whenever K changed (k_ver).
static int refresh_operator(sgeo_mesh* m) {
  hipStream_t s = m->ctx->stream;
  const size_t nk = (size_t)m->nelmt * 6561;
  const size_t np = (size_t)m->nelmt * SGEO_KPACK;
  if (m->kfmt == SGEO_KFMT_FULL && m->kprec == SGEO_K_F32) {
    if (!m->d_K32) TRY(dalloc(&m->d_K32, nk, &m->bytes));
    if (m->k32_ver != m->k_ver) {
      launch_round_f32(s, (int64_t)nk, m->d_K64, m->d_K32);
      m->k32_ver = m->k_ver;
    }
  } else if (m->kfmt == SGEO_KFMT_PACKED && m->kprec == SGEO_K_F64) {
    if (!m->d_KP64) TRY(dalloc(&m->d_KP64, np, &m->bytes));
    if (m->kp64_ver != m->k_ver) {
      launch_pack_upper_f64(s, m->nelmt, m->d_K64, m->d_KP64);
      m->kp64_ver = m->k_ver;
    }
  } else if (m->kfmt == SGEO_KFMT_PACKED && m->kprec == SGEO_K_F32) {
    if (!m->d_KP32) TRY(dalloc(&m->d_KP32, np, &m->bytes));
    if (m->kp32_ver != m->k_ver) {
      launch_pack_upper_f32(s, m->nelmt, m->d_K64, m->d_KP32);
      m->kp32_ver = m->k_ver;
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int sgeo_mesh_set_operator(sgeo_mesh* m, int kprec, int kfmt) {
  if (!m || m->kprec < 0) return fail(SGEO_EARG, "mesh has no stiffness");
  if ((kprec != SGEO_K_F64 && kprec != SGEO_K_F32) ||
      (kfmt != SGEO_KFMT_FULL && kfmt != SGEO_KFMT_PACKED))
    return fail(SGEO_EARG, "operator kprec=%d kfmt=%d", kprec, kfmt);
  HIPCHK(hipSetDevice(m->ctx->device));
  m->kprec = kprec;
  m->kfmt = kfmt;
  TRY(refresh_operator(m));
  HIPCHK(hipStreamSynchronize(m->ctx->stream));
  return 0;
}

extern "C" int sgeo_mesh_build_stiffness(sgeo_mesh* m, const double* gam,
                                         const double* ym, const double* nu,
                                         int kprec, double* extload,
                                         double* dprecon_out) {
  if (!m || !ym || !nu || !m->d_corners || !m->d_mat)
    return fail(SGEO_EARG, "sgeo_mesh_build_stiffness: mesh lacks corners/mat_id");
  if (extload && !gam) return fail(SGEO_EARG, "extload requested without gam");
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  TRY(set_k_storage(m, kprec));
  const int64_t n1 = (int64_t)m->neq + 1;
  double *d_ym = nullptr, *d_nu = nullptr, *d_gam = nullptr, *d_kdiag = nullptr,
         *d_eld = nullptr, *d_ext = nullptr;
  DevScratch tmp;
  TRY(tmp.alloc(&d_ym, m->nmat));
  TRY(tmp.alloc(&d_nu, m->nmat));
  TRY(tmp.alloc(&d_gam, m->nmat));
  TRY(tmp.alloc(&d_kdiag, (size_t)m->nelmt * 81));
  TRY(tmp.alloc(&d_eld, (size_t)m->nelmt * 81));
  TRY(tmp.alloc(&d_ext, (size_t)n1));
  HIPCHK(hipMemcpyAsync(d_ym, ym, m->nmat * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_nu, nu, m->nmat * 8, hipMemcpyHostToDevice, s));
  if (gam) HIPCHK(hipMemcpyAsync(d_gam, gam, m->nmat * 8, hipMemcpyHostToDevice, s));
  launch_stiffness(s, m->nelmt, m->d_corners, m->d_mat, d_ym, d_nu, m->d_K64,
                   d_kdiag, d_eld);
  // dprecon(egdof(k)) += km(k,k) in element order, then 1/d
  launch_assemble(s, n1, m->d_rowptr, m->d_slots, d_kdiag, m->d_dinv, c->d_sc, 0);
  launch_invert_diag(s, n1, m->d_dinv);
  if (extload) {
    // extload(egdof) = extload(egdof) - eld*gam, element order
    launch_scale_by_mat(s, m->nelmt, d_eld, m->d_mat, d_gam, d_kdiag);
    HIPCHK(hipMemsetAsync(d_ext, 0, n1 * 8, s));
    launch_assemble_acc(s, n1, m->d_rowptr, m->d_slots, d_kdiag, d_ext, 1, nullptr);
    HIPCHK(hipMemcpyAsync(extload, d_ext, n1 * 8, hipMemcpyDeviceToHost, s));
  }
  m->k_ver++;
  TRY(refresh_operator(m));
  if (dprecon_out)
    HIPCHK(hipMemcpyAsync(dprecon_out, m->d_dinv, n1 * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int sgeo_mesh_upload_stiffness(sgeo_mesh* m, const double* k,
                                          const double* dprecon, int kprec) {
  if (!m || !k || !dprecon) return fail(SGEO_EARG, "null argument");
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  TRY(set_k_storage(m, kprec));
  const size_t nk = (size_t)m->nelmt * 6561;
  HIPCHK(hipMemcpyAsync(m->d_K64, k, nk * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(m->d_dinv, dprecon, ((size_t)m->neq + 1) * 8,
                        hipMemcpyHostToDevice, s));
  m->k_ver++;
  TRY(refresh_operator(m));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int sgeo_mesh_download_stiffness(sgeo_mesh* m, double* k) {
  if (!m || !k || !m->d_K64) return fail(SGEO_EARG, "no FP64 stiffness");
  HIPCHK(hipSetDevice(m->ctx->device));
  HIPCHK(hipMemcpyAsync(k, m->d_K64, (size_t)m->nelmt * 6561 * 8,
                        hipMemcpyDeviceToHost, m->ctx->stream));
  HIPCHK(hipStreamSynchronize(m->ctx->stream));
  return 0;
}

static void ebe_elem(sgeo_mesh* m, const double* p, double* Y, int skip_mode) {
  hipStream_t s = m->ctx->stream;
  const PcgScalars* sc = m->ctx->d_sc;
  if (m->kfmt == SGEO_KFMT_PACKED) {
    if (m->kprec == SGEO_K_F32)
      launch_ebe_packed_f32(s, m->nelmt, m->d_KP32, m->d_gdof, p, Y, sc, skip_mode);
    else
      launch_ebe_packed_f64(s, m->nelmt, m->ctx->n_cu, m->d_KP64, m->d_gdof, p, Y, sc,
                            skip_mode);
  } else if (m->kprec == SGEO_K_F32) {
    launch_ebe_elem_f32k(s, m->nelmt, m->d_K32, m->d_gdof, p, Y, sc, skip_mode);
  } else {
    launch_ebe_elem_f64(s, m->nelmt, m->d_K64, m->d_gdof, p, Y, sc, skip_mode);
  }
}

// --------------------------------------------------------------------------
// ingest: GLL numbering on the device (src/mesh_spec.f90:17-386,
// src/apply_bc.f90:36-137)
// --------------------------------------------------------------------------
struct sgeo_spec {
  sgeo_ctx* ctx = nullptr;
  SpecDev d;
};

extern "C" int sgeo_hex2spec(sgeo_ctx* c, int32_t nelmt, int32_t nnode8,
                             const double* coord8, const int32_t* conn, int32_t nbc,
                             const int32_t* bc, sgeo_spec** out, int32_t* nnode,
                             int32_t* neq) {
  if (!c || !coord8 || !conn || !out || nelmt <= 0 || nnode8 <= 0 || nbc < 0 ||
      (nbc > 0 && !bc))
    return fail(SGEO_EARG, "sgeo_hex2spec: bad arguments");
  if ((int64_t)nelmt * 81 > INT32_MAX)
    return fail(SGEO_EARG, "sgeo_hex2spec: nelmt*81 exceeds int32");
  for (int64_t q = 0; q < (int64_t)nelmt * 8; q++)
    if (conn[q] < 1 || conn[q] > nnode8)
      return fail(SGEO_EARG, "sgeo_hex2spec: conn[%lld]=%d outside 1..%d", (long long)q,
                  conn[q], nnode8);
  for (int64_t r = 0; r < nbc; r++)
    if (bc[r * 3] < 1 || bc[r * 3] > nelmt || bc[r * 3 + 1] < 1 || bc[r * 3 + 1] > 6 ||
        bc[r * 3 + 2] < 0 || bc[r * 3 + 2] > 2)
      return fail(SGEO_EARG, "sgeo_hex2spec: boundary row %lld (%d, %d, %d) invalid",
                  (long long)r, bc[r * 3], bc[r * 3 + 1], bc[r * 3 + 2]);
  // tolerance of get_global: 1e-10 * |xmax - xmin| over the hex8 nodes
  // (src/mesh_spec.f90:178)
  double xmin = coord8[0], xmax = coord8[0];
  for (int64_t g = 1; g < nnode8; g++) {
    xmin = std::min(xmin, coord8[g * 3]);
    xmax = std::max(xmax, coord8[g * 3]);
  }
  const double tol = 1.e-10 * std::fabs(xmax - xmin);
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  double* d_coord = nullptr;
  int32_t *d_conn = nullptr, *d_bc = nullptr;
  DevScratch tmp;
  TRY(tmp.alloc(&d_coord, (size_t)nnode8 * 3));
  TRY(tmp.alloc(&d_conn, (size_t)nelmt * 8));
  TRY(tmp.alloc(&d_bc, (size_t)std::max(nbc, 1) * 3));
  HIPCHK(hipMemcpyAsync(d_coord, coord8, (size_t)nnode8 * 24, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_conn, conn, (size_t)nelmt * 32, hipMemcpyHostToDevice, s));
  if (nbc > 0)
    HIPCHK(hipMemcpyAsync(d_bc, bc, (size_t)nbc * 12, hipMemcpyHostToDevice, s));
  SpecDev d;
  const hipError_t e = hex2spec_device(s, nelmt, d_coord, d_conn, nbc, d_bc, tol, &d);
  if (e != hipSuccess)
    return fail(SGEO_EHIP, "sgeo_hex2spec: %s", hipGetErrorString(e));
  sgeo_spec* sp = new sgeo_spec();
  sp->ctx = c;
  sp->d = d;
  *out = sp;
  if (nnode) *nnode = d.nnode;
  if (neq) *neq = d.neq;
  return 0;
}

extern "C" int sgeo_spec_fetch(sgeo_spec* sp, int32_t* g_num, int32_t* gdof,
                               int32_t* gdof_elmt, double* g_coord, double* corners) {
  if (!sp) return fail(SGEO_EARG, "null spec");
  HIPCHK(hipSetDevice(sp->ctx->device));
  hipStream_t s = sp->ctx->stream;
  const SpecDev& d = sp->d;
  const size_t ne = (size_t)d.nelmt, nn = (size_t)d.nnode;
  if (g_num) HIPCHK(hipMemcpyAsync(g_num, d.g_num, ne * 27 * 4, hipMemcpyDeviceToHost, s));
  if (gdof) HIPCHK(hipMemcpyAsync(gdof, d.gdof, nn * 3 * 4, hipMemcpyDeviceToHost, s));
  if (gdof_elmt)
    HIPCHK(hipMemcpyAsync(gdof_elmt, d.gdof_elmt, ne * 81 * 4, hipMemcpyDeviceToHost, s));
  if (g_coord)
    HIPCHK(hipMemcpyAsync(g_coord, d.g_coord, nn * 3 * 8, hipMemcpyDeviceToHost, s));
  if (corners)
    HIPCHK(hipMemcpyAsync(corners, d.corners, ne * 24 * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int sgeo_spec_destroy(sgeo_spec* sp) {
  if (!sp) return 0;
  hipSetDevice(sp->ctx->device);
  hipStreamSynchronize(sp->ctx->stream);
  spec_device_free(&sp->d);
  delete sp;
  return 0;
}

extern "C" int sgeo_calibrate_copy(sgeo_ctx* c, size_t bytes, int reps,
                                   double* gbs) {
  if (!c || !gbs || bytes < 1024 || reps < 1)
    return fail(SGEO_EARG, "sgeo_calibrate_copy: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  const size_t n2 = bytes / 16;
  double *a = nullptr, *b = nullptr;
  DevScratch tmp;
  TRY(tmp.alloc(&a, n2 * 2));
  TRY(tmp.alloc(&b, n2 * 2));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipMemsetAsync(a, 0, n2 * 16, c->stream));
  launch_copy(c->stream, (int64_t)n2, a, b);  // warm-up
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int i = 0; i < reps; i++) launch_copy(c->stream, (int64_t)n2, a, b);
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *gbs = 2.0 * (double)n2 * 16.0 * reps / (ms * 1e-3) / 1e9;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return 0;
}

// Host arrays in, host arrays out: the device point functions of the
// plastic body (parity checks of the point math).
extern "C" int sgeo_constitutive(sgeo_ctx* c, int32_t n, const double* rec, double* out) {
  if (!c || !rec || !out || n <= 0) return fail(SGEO_EARG, "sgeo_constitutive: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  double *d_rec = nullptr, *d_out = nullptr;
  DevScratch tmp;
  TRY(tmp.alloc(&d_rec, (size_t)n * 9));
  TRY(tmp.alloc(&d_out, (size_t)n * 7));
  HIPCHK(hipMemcpyAsync(d_rec, rec, (size_t)n * 72, hipMemcpyHostToDevice, s));
  launch_constitutive(s, n, d_rec, d_out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, d_out, (size_t)n * 56, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int sgeo_math(sgeo_ctx* c, int32_t fn, int32_t n, const double* x, double* y) {
  if (!c || !x || !y || n <= 0 || fn < 0 || fn > 3)
    return fail(SGEO_EARG, "sgeo_math: bad arguments");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  double *d_x = nullptr, *d_y = nullptr;
  DevScratch tmp;
  TRY(tmp.alloc(&d_x, (size_t)n));
  TRY(tmp.alloc(&d_y, (size_t)n));
  HIPCHK(hipMemcpyAsync(d_x, x, (size_t)n * 8, hipMemcpyHostToDevice, s));
  launch_math(s, n, fn, d_x, d_y);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(y, d_y, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int sgeo_mesh_matvec(sgeo_mesh* m, const double* p, double* kp) {
  if (!m || !p || !kp || m->kprec < 0) return fail(SGEO_EARG, "no stiffness");
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int64_t n1 = (int64_t)m->neq + 1;
  HIPCHK(hipMemcpyAsync(m->d_p, p, n1 * 8, hipMemcpyHostToDevice, s));
  ebe_elem(m, m->d_p, m->d_Y, 0);
  launch_assemble(s, n1, m->d_rowptr, m->d_slots, m->d_Y, m->d_kp, c->d_sc, 0);
  HIPCHK(hipMemcpyAsync(kp, m->d_kp, n1 * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

// --------------------------------------------------------------------------
// PCG on device vectors (src/solver.f90:14-79)
// --------------------------------------------------------------------------
// Iterations are enqueued in chunks; each kernel of an iteration returns at
// once when the device-side stop flag is set, so over-enqueued iterations
// cost only their launches.  The host reads the flag once per chunk.
static int pcg_device(sgeo_mesh* m, double* d_u, const double* d_f, double tol,
                      int32_t maxiter, int32_t* cg_iter, int32_t* errcode,
                      double* part = nullptr) {
  sgeo_ctx* c = m->ctx;
  if (!part) part = c->d_part;
  hipStream_t s = c->stream;
  const int64_t n1 = (int64_t)m->neq + 1;
  if (maxiter < 1) {
    // the reference's DO loop body never runs: cg_iter = 1, errcode = -1
    *cg_iter = 1;
    *errcode = -1;
    return 0;
  }
  launch_pcg_begin(s, n1, d_u, part, c->d_sc, tol, maxiter);
  ebe_elem(m, d_u, m->d_Y, 2);
  launch_assemble(s, n1, m->d_rowptr, m->d_slots, m->d_Y, m->d_kp, c->d_sc, 2);
  const int seq = c->reduction == SGEO_REDUCE_SEQUENTIAL;
  launch_pcg_init(s, n1, d_f, m->d_kp, m->d_dinv, m->d_r, m->d_z, m->d_p,
                  part, c->d_sc, seq);
  int32_t enq = 0, iter_prev = 0;
  int chunk = 8;
  c->ev_used = 0;
  for (;;) {
    const int n = std::min<int32_t>(chunk, maxiter - enq);
    for (int it = 0; it < n; it++) {
      TRY(prof_mark(c));
      ebe_elem(m, m->d_p, m->d_Y, 1);
      TRY(prof_mark(c));
      launch_pcg_assemble_dot(s, n1, m->d_rowptr, m->d_slots, m->d_Y, m->d_p,
                              m->d_kp, part, c->d_sc, seq);
      launch_pcg_update(s, n1, m->d_p, m->d_kp, m->d_dinv, d_u, m->d_r,
                        m->d_z, part, c->d_sc, seq);
      launch_pcg_pupdate(s, n1, m->d_z, m->d_p, c->d_sc);
    }
    enq += n;
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_sc, c->d_sc, sizeof(PcgScalars),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (c->prof) {
      // only the launches of iterations that ran (later ones return at once)
      const int ran = c->h_sc->iter - iter_prev;
      c->ev_used = std::min(c->ev_used, 2 * std::max(ran, 0));
      TRY(prof_flush(c));
      iter_prev = c->h_sc->iter;
    }
    if (c->h_sc->done) break;
    if (enq >= maxiter)
      return fail(SGEO_ESTATE, "PCG: %d iterations enqueued without a verdict",
                  enq);
    chunk = std::min(chunk * 2, 64);
  }
  *cg_iter = c->h_sc->conv_iter;
  *errcode = c->h_sc->done == 1 ? 0 : -1;
  return 0;
}

// --------------------------------------------------------------------------
// Seam 1: sgeo_pcg_solver
// --------------------------------------------------------------------------
static uint64_t fp_mix(uint64_t h, const void* p, size_t nbytes, size_t nsamp) {
  const unsigned char* b = (const unsigned char*)p;
  const size_t step = std::max<size_t>(1, nbytes / 8 / nsamp) * 8;
  for (size_t o = 0; o + 8 <= nbytes; o += step) {
    uint64_t v;
    memcpy(&v, b + o, 8);
    h ^= v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
  }
  if (nbytes >= 8) {
    uint64_t v;
    memcpy(&v, b + nbytes - 8, 8);
    h ^= v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
  }
  return h;
}

static void set_errtag(char* errtag, const char* msg) {
  if (!errtag) return;
  size_t n = strlen(msg);
  if (n > 250) n = 250;
  memcpy(errtag, msg, n);
  memset(errtag + n, ' ', 250 - n);
}

// Drops the drop-in's cached device copies of k, dprecon and gdof_elmt.
extern "C" int sgeo_pcg_solver_invalidate(sgeo_ctx* c) {
  if (!c) return fail(SGEO_EARG, "null context");
  if (c->cache) {
    hipSetDevice(c->device);
    sgeo_mesh_destroy(c->cache);
  }
  c->cache = nullptr;
  c->cache_k = c->cache_g = c->cache_d = nullptr;
  c->cache_sig = 0;
  return 0;
}

extern "C" int sgeo_pcg_solver(sgeo_ctx* c, int32_t neq, int32_t nelmt,
                               const double* k, double* u, const double* f,
                               const double* dprecon, const int32_t* gdof,
                               double cg_tol, int32_t cg_maxiter,
                               int32_t* cg_iter, int32_t* errcode,
                               char* errtag) {
  if (!c || !k || !u || !f || !dprecon || !gdof || !cg_iter || !errcode ||
      neq <= 0 || nelmt <= 0)
    return fail(SGEO_EARG, "sgeo_pcg_solver: bad arguments");
  set_errtag(errtag, "ERROR: unknown!");
  *errcode = -1;
  HIPCHK(hipSetDevice(c->device));
  const size_t nk = (size_t)nelmt * 6561;
  const int64_t n1 = (int64_t)neq + 1;
  // Cache key: every value of dprecon and gdof_elmt (O(neq), a few ms on
  // the largest meshes) and 8192 strided values of k.  Any change of a
  // material's stiffness changes diag(K) and so dprecon (the reference's
  // phinu path rebuilds both in place, src/slope.F90:259-270).
  uint64_t sig = 1469598103934665603ULL ^ ((uint64_t)nelmt << 32) ^ (uint64_t)neq;
  sig = fp_mix(sig, k, nk * 8, 8192);
  sig = fp_mix(sig, dprecon, n1 * 8, (size_t)n1);
  sig = fp_mix(sig, gdof, (size_t)nelmt * 81 * 4, (size_t)nelmt * 81 / 2);
  if (!c->cache || c->cache_k != k || c->cache_g != gdof ||
      c->cache_d != dprecon || c->cache_sig != sig ||
      c->cache->nelmt != nelmt || c->cache->neq != neq) {
    TRY(sgeo_pcg_solver_invalidate(c));
    sgeo_mesh* nm = nullptr;
    TRY(sgeo_mesh_create(c, nelmt, neq, nullptr, gdof, nullptr, 0, &nm));
    Partial<sgeo_mesh, sgeo_mesh_destroy> guard{nm};
    TRY(sgeo_mesh_upload_stiffness(nm, k, dprecon, SGEO_K_F64));
    c->cache = guard.release();  // keyed only once the upload succeeded
    c->cache_k = k;
    c->cache_g = gdof;
    c->cache_d = dprecon;
    c->cache_sig = sig;
  }
  if ((size_t)n1 > c->uf_cap) {
    if (c->d_u) hipFree(c->d_u);
    if (c->d_f) hipFree(c->d_f);
    c->d_u = c->d_f = nullptr;
    c->uf_cap = 0;
    TRY(dalloc(&c->d_u, (size_t)n1));
    TRY(dalloc(&c->d_f, (size_t)n1));
    c->uf_cap = (size_t)n1;
  }
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(c->d_u, u, n1 * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->d_f, f, n1 * 8, hipMemcpyHostToDevice, s));
  TRY(pcg_device(c->cache, c->d_u, c->d_f, cg_tol, cg_maxiter, cg_iter, errcode));
  HIPCHK(hipMemcpyAsync(u, c->d_u, n1 * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (*errcode != 0) set_errtag(errtag, "ERROR: PCG solver doesn't converge!");
  else set_errtag(errtag, "");
  return 0;
}

// --------------------------------------------------------------------------
// Seam 1 without allocation: device arrays in, the inverse map and every
// work vector in the context's scratch array
// --------------------------------------------------------------------------
static inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of scratch one call of sgeo_pcg_solver_noalloc takes (rowptr, slots,
// element products, r, z, p, kp, the reduction partials and a flag word).
static size_t noalloc_bytes(int64_t neq, int64_t nelmt) {
  const int64_t n1 = neq + 1, ns = nelmt * 81;
  const size_t npart = 4 * (size_t)vec_reduce_blocks(n1) + 64;
  return align256((size_t)(n1 + 1) * 4) + align256((size_t)ns * 4) +
         align256((size_t)ns * 8) + 4 * align256((size_t)n1 * 8) + align256(npart * 8) +
         align256(4);
}

extern "C" size_t sgeo_pcg_solver_noalloc_bytes(int32_t neq, int32_t nelmt) {
  if (neq <= 0 || nelmt <= 0) return 0;
  return noalloc_bytes(neq, nelmt);
}

extern "C" int sgeo_pcg_solver_noalloc(sgeo_ctx* c, int32_t neq, int32_t nelmt,
                                       const double* d_k, double* d_u, const double* d_f,
                                       const double* d_dprecon, const int32_t* d_gdof,
                                       double cg_tol, int32_t cg_maxiter, int32_t* cg_iter,
                                       int32_t* errcode, char* errtag) {
  if (!c || !d_k || !d_u || !d_f || !d_dprecon || !d_gdof || !cg_iter || !errcode ||
      neq <= 0 || nelmt <= 0)
    return fail(SGEO_EARG, "sgeo_pcg_solver_noalloc: bad arguments");
  if ((int64_t)nelmt * 81 > INT32_MAX)
    return fail(SGEO_EARG, "sgeo_pcg_solver_noalloc: nelmt*81 exceeds int32");
  const size_t need = noalloc_bytes(neq, nelmt);
  if (need > (size_t)SGEO_NOALLOC_SCRATCH_BYTES)
    return fail(SGEO_EARG,
                "sgeo_pcg_solver_noalloc: neq=%d, nelmt=%d need %zu bytes of scratch, "
                "more than its %zu",
                neq, nelmt, need, (size_t)SGEO_NOALLOC_SCRATCH_BYTES);
  if (!c->d_scratch) return fail(SGEO_ESTATE, "sgeo_pcg_solver_noalloc: no scratch");
  set_errtag(errtag, "ERROR: unknown!");
  *errcode = -1;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int64_t n1 = (int64_t)neq + 1, ns = (int64_t)nelmt * 81;
  // carve the scratch: each piece 256-byte aligned
  unsigned char* q = c->d_scratch;
  auto take = [&q](size_t bytes) {
    unsigned char* r = q;
    q += align256(bytes);
    return r;
  };
  sgeo_mesh m;  // a view that owns nothing: the caller's K, dprecon and gdof,
                // work arrays in the scratch
  m.ctx = c;
  m.nelmt = nelmt;
  m.neq = neq;
  m.kprec = SGEO_K_F64;
  m.kfmt = SGEO_KFMT_FULL;
  m.d_K64 = const_cast<double*>(d_k);
  m.d_dinv = const_cast<double*>(d_dprecon);
  m.d_gdof = const_cast<int32_t*>(d_gdof);
  m.d_rowptr = reinterpret_cast<int32_t*>(take((size_t)(n1 + 1) * 4));
  m.d_slots = reinterpret_cast<int32_t*>(take((size_t)ns * 4));
  m.d_Y = reinterpret_cast<double*>(take((size_t)ns * 8));
  m.d_r = reinterpret_cast<double*>(take((size_t)n1 * 8));
  m.d_z = reinterpret_cast<double*>(take((size_t)n1 * 8));
  m.d_p = reinterpret_cast<double*>(take((size_t)n1 * 8));
  m.d_kp = reinterpret_cast<double*>(take((size_t)n1 * 8));
  double* part = reinterpret_cast<double*>(take((4 * (size_t)vec_reduce_blocks(n1) + 64) * 8));
  int32_t* bad = reinterpret_cast<int32_t*>(take(4));
  // the equation -> slot map of gdof_elmt (z is the fill's cursor), then a
  // check of the equation range before any product reads through gdof
  launch_csr_build(s, ns, neq, d_gdof, m.d_rowptr, m.d_slots,
                   reinterpret_cast<int32_t*>(m.d_z), bad);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->h_flag, bad, 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (*c->h_flag)
    return fail(SGEO_EARG, "sgeo_pcg_solver_noalloc: gdof_elmt entry outside 0..%d", neq);
  TRY(pcg_device(&m, d_u, d_f, cg_tol, cg_maxiter, cg_iter, errcode, part));
  if (*errcode != 0) set_errtag(errtag, "ERROR: PCG solver doesn't converge!");
  else set_errtag(errtag, "");
  return 0;
}

// --------------------------------------------------------------------------
// Seam 2: slope plastic iteration
// --------------------------------------------------------------------------
extern "C" int sgeo_slope_create(sgeo_mesh* m, const double* extload,
                                 sgeo_slope** out) {
  if (!m || !extload || !out || m->kprec < 0 || !m->d_corners || !m->d_mat)
    return fail(SGEO_EARG, "sgeo_slope_create: mesh needs stiffness, corners, mat_id");
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  sgeo_slope* t = new sgeo_slope();
  Partial<sgeo_slope, sgeo_slope_destroy> guard{t};
  t->mesh = m;
  const size_t n1 = (size_t)m->neq + 1;
  const size_t ns = (size_t)m->nelmt * 27 * 6;
  TRY(dalloc(&t->d_x, n1));
  TRY(dalloc(&t->d_oldx, n1));
  TRY(dalloc(&t->d_load, n1));
  TRY(dalloc(&t->d_body, n1));
  TRY(dalloc(&t->d_ext, n1));
  TRY(dalloc(&t->d_evpt, ns));
  TRY(dalloc(&t->d_stress, ns));
  TRY(dalloc(&t->d_BL, (size_t)m->nelmt * 81));
  TRY(dalloc(&t->d_flags, (size_t)m->nelmt * 27));
  TRY(dalloc(&t->d_mp, (size_t)std::max(1, m->nmat)));
  HIPCHK(hipMemcpyAsync(t->d_ext, extload, n1 * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *out = guard.release();
  return 0;
}

extern "C" int sgeo_slope_destroy(sgeo_slope* t) {
  if (!t) return 0;
  hipSetDevice(t->mesh->ctx->device);
  hipStreamSynchronize(t->mesh->ctx->stream);
  void* ptrs[] = {t->d_x,      t->d_oldx, t->d_load,  t->d_body,
                  t->d_ext,    t->d_evpt, t->d_stress, t->d_BL,
                  t->d_flags,  t->d_mp,   t->d_vm,     t->d_sf,
                  t->d_nacc};
  for (void* p : ptrs)
    if (p) hipFree(p);
  delete t;
  return 0;
}

extern "C" int sgeo_slope_begin_srf(sgeo_slope* t, const double* ym,
                                    const double* nuf, const double* phif,
                                    const double* cohf, const double* psif,
                                    double dt) {
  if (!t || !ym || !nuf || !phif || !cohf || !psif)
    return fail(SGEO_EARG, "sgeo_slope_begin_srf: null argument");
  sgeo_mesh* m = t->mesh;
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  t->h_mp.resize((size_t)m->nmat);
  for (int i = 0; i < m->nmat; i++)
    t->h_mp[i] = MatParams{ym[i], nuf[i], phif[i], cohf[i], psif[i], dt};
  const size_t n1 = (size_t)m->neq + 1;
  const size_t ns = (size_t)m->nelmt * 27 * 6;
  HIPCHK(hipMemcpyAsync(t->d_mp, t->h_mp.data(), m->nmat * sizeof(MatParams),
                        hipMemcpyHostToDevice, s));
  // stress_elmt=0 (:249); bodyload=0, evpt=0, x=0, oldx=0 (:281-282)
  HIPCHK(hipMemsetAsync(t->d_stress, 0, ns * 8, s));
  HIPCHK(hipMemsetAsync(t->d_evpt, 0, ns * 8, s));
  HIPCHK(hipMemsetAsync(t->d_body, 0, n1 * 8, s));
  HIPCHK(hipMemsetAsync(t->d_x, 0, n1 * 8, s));
  HIPCHK(hipMemsetAsync(t->d_oldx, 0, n1 * 8, s));
  HIPCHK(hipMemsetAsync(t->d_flags, 0, (size_t)m->nelmt * 27, s));
  t->begun = true;
  return 0;
}

// Load increment of an excavation stage (src/excavation.F90:509-512):
// bodyload=0, evpt=0, x=0, oldx=0, the stress state kept (stress0 != NULL
// replaces it, reference layout (6,27,nelmt)); the body then also records
// the von Mises plastic strain at the final pass (src/excavation.F90:601-602).
extern "C" int sgeo_slope_begin_increment(sgeo_slope* t, const double* ym,
                                          const double* nuf, const double* phif,
                                          const double* cohf, const double* psif,
                                          double dt, const double* stress0) {
  if (!t || !ym || !nuf || !phif || !cohf || !psif)
    return fail(SGEO_EARG, "sgeo_slope_begin_increment: null argument");
  sgeo_mesh* m = t->mesh;
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  t->h_mp.resize((size_t)m->nmat);
  for (int i = 0; i < m->nmat; i++)
    t->h_mp[i] = MatParams{ym[i], nuf[i], phif[i], cohf[i], psif[i], dt};
  const size_t n1 = (size_t)m->neq + 1;
  const size_t ns = (size_t)m->nelmt * 27 * 6;
  if (!t->d_vm) TRY(dalloc(&t->d_vm, (size_t)m->nelmt * 27));
  if (!t->d_sf) TRY(dalloc(&t->d_sf, (size_t)m->nelmt * 27));
  HIPCHK(hipMemcpyAsync(t->d_mp, t->h_mp.data(), m->nmat * sizeof(MatParams),
                        hipMemcpyHostToDevice, s));
  if (stress0)
    HIPCHK(hipMemcpyAsync(t->d_stress, stress0, ns * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(t->d_evpt, 0, ns * 8, s));
  HIPCHK(hipMemsetAsync(t->d_body, 0, n1 * 8, s));
  HIPCHK(hipMemsetAsync(t->d_x, 0, n1 * 8, s));
  HIPCHK(hipMemsetAsync(t->d_oldx, 0, n1 * 8, s));
  HIPCHK(hipMemsetAsync(t->d_flags, 0, (size_t)m->nelmt * 27, s));
  HIPCHK(hipMemsetAsync(t->d_vm, 0, (size_t)m->nelmt * 27 * 8, s));
  HIPCHK(hipMemsetAsync(t->d_sf, 0, (size_t)m->nelmt * 27 * 8, s));
  HIPCHK(hipStreamSynchronize(s));
  t->excav = 1;
  t->begun = true;
  return 0;
}

extern "C" int sgeo_slope_download_vmeps(sgeo_slope* t, double* vm_pt) {
  if (!t || !vm_pt || !t->d_vm) return fail(SGEO_ESTATE, "no excavation state");
  sgeo_mesh* m = t->mesh;
  HIPCHK(hipSetDevice(m->ctx->device));
  HIPCHK(hipMemcpyAsync(vm_pt, t->d_vm, (size_t)m->nelmt * 27 * 8,
                        hipMemcpyDeviceToHost, m->ctx->stream));
  HIPCHK(hipStreamSynchronize(m->ctx->stream));
  return 0;
}

// pcg_solver on a resident mesh (its operator, its inverted diagonal):
// u (host, neq+1, warm start, overwritten), f (host, neq+1).
extern "C" int sgeo_mesh_solve(sgeo_mesh* m, double* u, const double* f,
                               double cg_tol, int32_t cg_maxiter, int32_t* cg_iter,
                               int32_t* errcode) {
  if (!m || !u || !f || !cg_iter || !errcode || m->kprec < 0)
    return fail(SGEO_EARG, "sgeo_mesh_solve: bad arguments");
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const size_t n1 = (size_t)m->neq + 1;
  double *d_u = nullptr, *d_f = nullptr;
  DevScratch tmp;
  TRY(tmp.alloc(&d_u, n1));
  TRY(tmp.alloc(&d_f, n1));
  HIPCHK(hipMemcpyAsync(d_u, u, n1 * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_f, f, n1 * 8, hipMemcpyHostToDevice, s));
  const int rc = pcg_device(m, d_u, d_f, cg_tol, cg_maxiter, cg_iter, errcode);
  if (rc == 0) {
    HIPCHK(hipMemcpyAsync(u, d_u, n1 * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return rc;
}

// Elastic stress at every GLL point from a displacement x (host, neq+1):
// src/postprocess.f90:40-79 with E = ym[m], nu = nu[m].  stress: host,
// (6,27,nelmt) reference layout.
extern "C" int sgeo_mesh_elastic_stress(sgeo_mesh* m, const double* ym,
                                        const double* nu, const double* x,
                                        double* stress) {
  if (!m || !ym || !nu || !x || !stress || !m->d_corners || !m->d_mat)
    return fail(SGEO_EARG, "sgeo_mesh_elastic_stress: bad arguments");
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const size_t n1 = (size_t)m->neq + 1, ns = (size_t)m->nelmt * 27 * 6;
  const int nmat = std::max(1, m->nmat);
  double *d_x = nullptr, *d_st = nullptr, *d_mat = nullptr;
  DevScratch tmp;
  TRY(tmp.alloc(&d_x, n1));
  TRY(tmp.alloc(&d_st, ns));
  TRY(tmp.alloc(&d_mat, 2 * (size_t)nmat));
  HIPCHK(hipMemcpyAsync(d_x, x, n1 * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_mat, ym, nmat * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_mat + nmat, nu, nmat * 8, hipMemcpyHostToDevice, s));
  launch_elastic_stress(s, m->nelmt, m->d_corners, m->d_gdof, m->d_mat, d_mat,
                        d_mat + nmat, d_x, d_st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(stress, d_st, ns * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

// Excavation load of the mesh's elements -- the removed ones of a stage,
// numbered with the stage's equations (src/excavation_library.f90:186-236
// and distribute2ghosts, src/serial_library.f90:110-130): load (host,
// neq+1) = per equation the sum of its element contributions in element
// order, load[0] = 0.  stress: host, (6,27,nelmt) of these elements.
extern "C" int sgeo_mesh_excavation_load(sgeo_mesh* m, const double* gam,
                                         const double* stress, double* load) {
  if (!m || !gam || !stress || !load || !m->d_corners || !m->d_mat)
    return fail(SGEO_EARG, "sgeo_mesh_excavation_load: bad arguments");
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int64_t n1 = (int64_t)m->neq + 1;
  const size_t ns = (size_t)m->nelmt * 27 * 6;
  const int nmat = std::max(1, m->nmat);
  double *d_st = nullptr, *d_gam = nullptr, *d_out = nullptr;
  DevScratch tmp;
  TRY(tmp.alloc(&d_st, ns));
  TRY(tmp.alloc(&d_gam, (size_t)nmat));
  TRY(tmp.alloc(&d_out, (size_t)n1));
  HIPCHK(hipMemcpyAsync(d_st, stress, ns * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_gam, gam, nmat * 8, hipMemcpyHostToDevice, s));
  launch_excav_load(s, m->nelmt, m->d_corners, m->d_mat, d_gam, d_st, m->d_Y);
  launch_assemble(s, n1, m->d_rowptr, m->d_slots, m->d_Y, d_out, nullptr, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(load, d_out, n1 * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int sgeo_slope_plastic_iter(sgeo_slope* t, double cg_tol,
                                       int32_t cg_maxiter, double nl_tol,
                                       int is_last, int32_t* cg_iter,
                                       int32_t* errcode, double* uerr,
                                       int32_t* nl_isconv, double* fmax) {
  if (!t || !t->begun) return fail(SGEO_ESTATE, "plastic_iter before begin_srf");
  sgeo_mesh* m = t->mesh;
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int64_t n1 = (int64_t)m->neq + 1;
  int32_t it = 0, ec = -1;
  launch_load(s, n1, t->d_ext, t->d_body, t->d_load);
  TRY(pcg_device(m, t->d_x, t->d_load, cg_tol, cg_maxiter, &it, &ec));
  if (cg_iter) *cg_iter = it;
  if (errcode) *errcode = ec;
  if (ec != 0) return 0;  // the caller stops (src/slope.F90:295)
  launch_uerr(s, n1, t->d_x, t->d_oldx, c->d_part, c->d_ps, nl_tol);
  BodyArgs b;
  b.nel = m->nelmt;
  b.corners = m->d_corners;
  b.gdof = m->d_gdof;
  b.mat_id = m->d_mat;
  b.mp = t->d_mp;
  b.nmat = m->nmat;
  b.x = t->d_x;
  b.evpt = t->d_evpt;
  b.stress = t->d_stress;
  b.BL = t->d_BL;
  b.flags = t->d_flags;
  b.ps = c->d_ps;
  b.is_last = is_last;
  b.excav = t->excav;
  b.vm_pt = t->d_vm;
  b.sf_pt = t->d_sf;
  launch_plastic_body(s, b);
  // bodyload(egdof) += bload unless converged/last (src/slope.F90:380-384);
  // the kernel itself returns when nl_isconv is set on the device
  if (!is_last)
    launch_assemble_acc(s, n1, m->d_rowptr, m->d_slots, t->d_BL, t->d_body, 0,
                        c->d_ps);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->h_ps, c->d_ps, sizeof(PlasticScalars),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const int conv = c->h_ps->nl_isconv;
  if (uerr) *uerr = c->h_ps->uerr;
  if (nl_isconv) *nl_isconv = conv;
  if (fmax) {
    unsigned long long b = c->h_ps->fmax_bits;
    double v;
    memcpy(&v, &b, 8);
    *fmax = v;
  }
  return 0;
}

extern "C" int sgeo_slope_run_srf(sgeo_slope* t, const int32_t* gdof_node,
                                  int32_t nnode, double cg_tol,
                                  int32_t cg_maxiter, double nl_tol,
                                  int32_t nl_maxiter, int32_t* cg_tot,
                                  int32_t* nl_tot, double* uxmax, double* umax,
                                  double* fmax, int32_t* errcode) {
  if (!t || !gdof_node || nl_maxiter < 1 || nnode < 0 || !cg_tot || !nl_tot || !uxmax ||
      !umax || !fmax || !errcode)
    return fail(SGEO_EARG, "sgeo_slope_run_srf: bad arguments");
  sgeo_mesh* m = t->mesh;
  for (int64_t q = 0; q < 3 * (int64_t)nnode; q++)
    if (gdof_node[q] < 0 || gdof_node[q] > m->neq)
      return fail(SGEO_EARG, "sgeo_slope_run_srf: gdof_node[%lld]=%d outside 0..%d",
                  (long long)q, gdof_node[q], m->neq);
  int32_t tot = 0, nl = 0, ec = 0;
  double fm = 0.0;
  for (nl = 1; nl <= nl_maxiter; nl++) {
    int32_t it = 0, conv = 0;
    double ue = 0.0;
    const int last = nl == nl_maxiter;
    TRY(sgeo_slope_plastic_iter(t, cg_tol, cg_maxiter, nl_tol, last, &it, &ec,
                                &ue, &conv, &fm));
    if (ec != 0) break;
    tot += it;
    if (conv || last) break;
  }
  if (nl > nl_maxiter) nl = nl_maxiter;
  *errcode = ec;
  *cg_tot = tot;
  *nl_tot = nl;
  *fmax = fm;
  // uxmax/umax over nodal displacements (src/slope.F90:401-413)
  std::vector<double> x((size_t)m->neq + 1);
  HIPCHK(hipMemcpyAsync(x.data(), t->d_x, x.size() * 8, hipMemcpyDeviceToHost,
                        m->ctx->stream));
  HIPCHK(hipStreamSynchronize(m->ctx->stream));
  double ux = 0.0, um = 0.0;
  for (int32_t n = 0; n < nnode; n++) {
    double u3[3];
    for (int cc = 0; cc < 3; cc++) {
      const int32_t q = gdof_node[3 * (size_t)n + cc];
      u3[cc] = q ? x[(size_t)q] : 0.0;
      ux = std::max(ux, std::fabs(u3[cc]));
    }
    um = std::max(um, std::sqrt(u3[0] * u3[0] + u3[1] * u3[1] + u3[2] * u3[2]));
  }
  *uxmax = ux;
  *umax = um;
  return 0;
}

extern "C" int sgeo_slope_download(sgeo_slope* t, double* x, double* bodyload,
                                   double* evpt, double* stress,
                                   uint8_t* flags) {
  if (!t) return fail(SGEO_EARG, "null slope");
  sgeo_mesh* m = t->mesh;
  hipStream_t s = m->ctx->stream;
  HIPCHK(hipSetDevice(m->ctx->device));
  const size_t n1 = (size_t)m->neq + 1;
  const size_t ns = (size_t)m->nelmt * 27 * 6;
  if (x) HIPCHK(hipMemcpyAsync(x, t->d_x, n1 * 8, hipMemcpyDeviceToHost, s));
  if (bodyload)
    HIPCHK(hipMemcpyAsync(bodyload, t->d_body, n1 * 8, hipMemcpyDeviceToHost, s));
  if (evpt) HIPCHK(hipMemcpyAsync(evpt, t->d_evpt, ns * 8, hipMemcpyDeviceToHost, s));
  if (stress)
    HIPCHK(hipMemcpyAsync(stress, t->d_stress, ns * 8, hipMemcpyDeviceToHost, s));
  if (flags)
    HIPCHK(hipMemcpyAsync(flags, t->d_flags, (size_t)m->nelmt * 27,
                          hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int sgeo_slope_accumulate_nodal(sgeo_slope* t, int32_t first) {
  if (!t || !t->excav || !t->d_vm || !t->d_sf)
    return fail(SGEO_ESTATE, "sgeo_slope_accumulate_nodal: no excavation increment");
  sgeo_mesh* m = t->mesh;
  if (!m->d_nrowptr || !m->d_nslots || m->nnode <= 0)
    return fail(SGEO_ESTATE, "no node map: call sgeo_mesh_set_nodes first");
  HIPCHK(hipSetDevice(m->ctx->device));
  if (t->d_nacc && t->nacc_nnode != m->nnode) {  // node map replaced since
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    hipFree(t->d_nacc);
    t->d_nacc = nullptr;
  }
  if (!t->d_nacc) {
    t->nacc_init = 0;
    TRY(dalloc(&t->d_nacc, 2 * (size_t)m->nnode));
    t->nacc_nnode = m->nnode;
  }
  if (!first && !t->nacc_init)
    return fail(SGEO_ESTATE,
                "sgeo_slope_accumulate_nodal: first=0 before an accumulation with first=1");
  launch_nodal_accumulate(m->ctx->stream, m->nnode, m->d_nrowptr, m->d_nslots, t->d_vm,
                          t->d_sf, t->d_nacc, first ? 1 : 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(m->ctx->stream));
  t->nacc_init = 1;
  return 0;
}

extern "C" int sgeo_slope_nodal_fields(sgeo_slope* t, float* stress, float* psigma,
                                       float* nsigma, float* scf, float* vmeps) {
  if (!t) return fail(SGEO_EARG, "null slope");
  const bool acc = t->excav && t->nacc_init && t->nacc_nnode == t->mesh->nnode;
  return nodal_fields(t->mesh, t->d_stress, acc ? t->d_nacc : nullptr, stress, psigma,
                      nsigma, scf, vmeps);
}

// --------------------------------------------------------------------------
// multi-trial mixed-precision engine
// --------------------------------------------------------------------------
// Host bookkeeping per column: a trial is a sequence of plastic iterations
// (src/slope.F90:285-392); each iteration is a PCG call (device state
// machine, BColState) followed by the element x GLL body.  Columns advance
// independently: a column whose PCG finished gets its body pass at the next
// host synchronisation and starts its next PCG call, while the others keep
// iterating.
enum BatchPhase { BPH_IDLE = 0, BPH_RUN = 1, BPH_ENDED = 2 };

#define STAGE_SLOTS_N 5  // pinned staging slots of a batch (see stage_upload)

struct sgeo_batch {
  sgeo_mesh* mesh = nullptr;
  int B = 0;
  int64_t n1 = 0;
  double *d_x = nullptr, *d_oldx = nullptr, *d_load = nullptr, *d_body = nullptr;
  double *d_Y64 = nullptr, *d_ext = nullptr, *d_xel = nullptr;
  float *d_r = nullptr, *d_p = nullptr;
  float *d_Y32 = nullptr, *d_dinv32 = nullptr;
  double *d_evpt = nullptr, *d_stress = nullptr, *d_BL = nullptr;
  uint8_t* d_flags = nullptr;
  MatParams* d_mp = nullptr;
  PlasticScalars *d_ps = nullptr, *h_ps = nullptr;
  BCol *d_bc = nullptr, *h_bc = nullptr;
  int32_t *d_cst = nullptr, *h_cst = nullptr;
  int32_t *d_mask = nullptr, *d_islast = nullptr, *d_cols = nullptr;
  // residual list of a residual step: columns [0..64), count at [64]; and
  // each listed column's place in it
  int32_t *d_rlist = nullptr, *d_rslot = nullptr;
  int32_t* d_eorder = nullptr;  // the MFMA product's walk over the elements
  double* d_part = nullptr;  // equation-chunk partials and their groups
  double* d_edot = nullptr;  // the elements' p.Kp shares and their groups
  // per column
  std::vector<int32_t> phase, nl_iter, cg_tot, need_start, seeded, seed_cg, err;
  std::vector<int32_t> last_it, last_rep;  // counters at the last synchronisation
  std::vector<int32_t> cst;  // host copy of the device states
  std::vector<double> fmax, uerr;
  std::vector<MatParams> h_mp;
  // elastic memo: the first PCG call of every trial solves K x = extload
  // from x = 0 (bodyload is zero), whatever the trial's strengths
  bool memo_valid = false;
  int32_t memo_cg = 0;
  int64_t stat_inner = 0, stat_outer = 0, stat_steps = 0;
  int k_ver = -1;  // mesh stiffness version the FP32 diagonal was made from
  size_t bytes = 0;
  // HIP graphs of 8, 16 and 32 steps (every launch argument is fixed for
  // the engine's lifetime: the steps read their state from the device)
  bool use_graphs = true;
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
  BmvGeom geom;  // launch geometry of the MFMA product
  // FP64 residuals (starting and replacing columns) are formed on the steps
  // gstep with gstep % resid_period == 0; gstep counts every step the
  // engine has run, so a column's residual steps fall at fixed counts of its
  // own iterations
  int resid_period = 1;
  int64_t gstep = 0;
  // Deferred x: a column starts on a residual step and iterates once on
  // every step until it ends, so its iteration i runs on a step with
  // gstep + 1 = i (mod resid_period), the same phase for every column.  x
  // is then written on the steps with (gstep + 1) % xlen == 0 only (xlen =
  // batch_xlen, a divisor of the period), and is current on every residual
  // step.  pring search-direction buffers hold the iterations not yet in x.
  bool defer_x = true;
  int pring = 2;  // p buffers: 8, or 2 when 8 n1 B reaches 2^32
  int32_t* h_stage = nullptr;  // pinned staging of the small uploads
  unsigned stage_busy = 0;
  double step_us = 0.0;  // running mean of the wall time of one step
  // excavation increments (sgeo_batch_begin_increment): per-point plastic
  // strain and stress concentration factor [col][e][27], and per-column
  // nodal accumulators [col][2][nnode]
  double *d_vm = nullptr, *d_sf = nullptr, *d_nacc = nullptr;
  int32_t nacc_nnode = 0;            // node count d_nacc was sized for
  std::vector<int32_t> nacc_valid;   // column's accumulators describe its state
  std::vector<int32_t> nacc_init;    // column's accumulators hold a started sum
};

static int ensure_packed(sgeo_mesh* m, int kprec) {
  const int kp = m->kprec, kf = m->kfmt;
  m->kprec = kprec;
  m->kfmt = SGEO_KFMT_PACKED;
  const int rc = refresh_operator(m);
  m->kprec = kp;
  m->kfmt = kf;
  return rc;
}

// Packed FP32/FP64 operators and the FP32 inverse diagonal follow the mesh's
// current stiffness (rebuilt when K changed since the last call).
static int batch_sync_operator(sgeo_batch* b) {
  sgeo_mesh* m = b->mesh;
  TRY(ensure_packed(m, SGEO_K_F32));
  TRY(ensure_packed(m, SGEO_K_F64));
  if (b->k_ver != m->k_ver) {
    launch_to_f32(m->ctx->stream, b->n1, m->d_dinv, b->d_dinv32);
    b->k_ver = m->k_ver;
    b->memo_valid = false;
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// The order in which the MFMA product walks the elements: by the Morton
// code of the element centroid on a grid of the same spacing on every axis,
// so the elements a product has in flight together (and each XCD's
// contiguous share of the walk) form compact regions whose gathered P rows
// meet in L2 and the MALL.  Y and the p.Kp shares stay indexed by element,
// so the order changes no result bit.
static int batch_walk_order(sgeo_mesh* m, sgeo_batch* b) {
  const int32_t ne = m->nelmt;
  std::vector<double> cor((size_t)ne * 24);
  HIPCHK(hipMemcpy(cor.data(), m->d_corners, cor.size() * 8, hipMemcpyDeviceToHost));
  std::vector<double> cen((size_t)ne * 3, 0.0);
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int32_t e = 0; e < ne; e++)
    for (int d = 0; d < 3; d++) {
      double c = 0.0;
      for (int k = 0; k < 8; k++) c += cor[(size_t)e * 24 + k * 3 + d];
      c *= 0.125;
      cen[(size_t)e * 3 + d] = c;
      lo[d] = std::min(lo[d], c);
      hi[d] = std::max(hi[d], c);
    }
  double ext = 0.0;
  for (int d = 0; d < 3; d++) ext = std::max(ext, hi[d] - lo[d]);
  const double scale = ext > 0.0 ? 2097151.0 / ext : 0.0;  // 21 bits per axis
  auto spread = [](uint64_t v) {  // bit i -> bit 3 i
    uint64_t r = 0;
    for (int i = 0; i < 21; i++) r |= ((v >> i) & 1ull) << (3 * i);
    return r;
  };
  std::vector<std::pair<uint64_t, int32_t>> key((size_t)ne);
  for (int32_t e = 0; e < ne; e++) {
    uint64_t code = 0;
    for (int d = 0; d < 3; d++) {
      const double q = std::floor((cen[(size_t)e * 3 + d] - lo[d]) * scale);
      const uint64_t u = (uint64_t)std::min(2097151.0, std::max(0.0, q));
      code |= spread(u) << d;
    }
    key[(size_t)e] = {code, e};
  }
  std::sort(key.begin(), key.end());
  std::vector<int32_t> order((size_t)ne);
  for (int32_t i = 0; i < ne; i++) order[(size_t)i] = key[(size_t)i].second;
  TRY(dalloc(&b->d_eorder, (size_t)ne, &b->bytes));
  HIPCHK(hipMemcpy(b->d_eorder, order.data(), (size_t)ne * 4, hipMemcpyHostToDevice));
  return 0;
}

extern "C" int sgeo_batch_create(sgeo_mesh* m, int32_t ncol, const double* extload,
                                 sgeo_batch** out) {
  if (!m || !extload || !out || m->kprec < 0 || !m->d_corners || !m->d_mat)
    return fail(SGEO_EARG, "sgeo_batch_create: mesh needs stiffness, corners, mat_id");
  if (ncol < 1 || ncol > 64 || (ncol & (ncol - 1)))
    return fail(SGEO_EARG, "sgeo_batch_create: ncol %d not a power of 2 <= 64", ncol);
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  sgeo_batch* b = new sgeo_batch();
  Partial<sgeo_batch, sgeo_batch_destroy> guard{b};
  b->mesh = m;
  b->B = ncol;
  b->n1 = (int64_t)m->neq + 1;
  const size_t nv = (size_t)b->n1 * ncol;
  // the element products address the search-direction buffers with 32-bit
  // offsets (pring n1 B entries, pring 8 or at least 2)
  if (2 * (uint64_t)nv >= (1ull << 32))
    return fail(SGEO_EARG, "sgeo_batch_create: %lld equations x %d columns too large",
                (long long)b->n1, ncol);
  b->pring = 8 * (uint64_t)nv < (1ull << 32) ? 8 : 2;
  const size_t ny = (size_t)m->nelmt * 81 * ncol;
  const size_t ns = (size_t)m->nelmt * 27 * 6 * ncol;
  size_t* acc = &b->bytes;
  TRY(dalloc(&b->d_x, nv, acc));
  TRY(dalloc(&b->d_oldx, nv, acc));
  TRY(dalloc(&b->d_load, nv, acc));
  TRY(dalloc(&b->d_body, nv, acc));
  TRY(dalloc(&b->d_Y64, ny, acc));
  TRY(dalloc(&b->d_ext, (size_t)b->n1, acc));
  TRY(dalloc(&b->d_xel, (size_t)b->n1, acc));
  TRY(dalloc(&b->d_r, nv, acc));
  TRY(dalloc(&b->d_p, (size_t)b->pring * nv, acc));  // search directions, i & (pring - 1)
  TRY(dalloc(&b->d_Y32, ny, acc));
  TRY(dalloc(&b->d_dinv32, (size_t)b->n1, acc));
  TRY(dalloc(&b->d_evpt, ns, acc));
  TRY(dalloc(&b->d_stress, ns, acc));
  TRY(dalloc(&b->d_BL, ny, acc));
  TRY(dalloc(&b->d_flags, (size_t)m->nelmt * 27 * ncol, acc));
  TRY(dalloc(&b->d_mp, (size_t)std::max(1, m->nmat) * ncol, acc));
  TRY(dalloc(&b->d_ps, (size_t)ncol, acc));
  TRY(dalloc(&b->d_bc, (size_t)ncol, acc));
  TRY(dalloc(&b->d_cst, (size_t)ncol, acc));
  TRY(dalloc(&b->d_mask, (size_t)ncol, acc));
  TRY(dalloc(&b->d_islast, (size_t)ncol, acc));
  TRY(dalloc(&b->d_cols, (size_t)ncol, acc));
  TRY(dalloc(&b->d_rlist, (size_t)65, acc));
  TRY(dalloc(&b->d_rslot, (size_t)64, acc));
  TRY(dalloc(&b->d_part, (size_t)batch_partials(b->n1) * ncol, acc));
  TRY(dalloc(&b->d_edot, (size_t)batch_element_partials(m->nelmt) * ncol, acc));
  TRY(batch_walk_order(m, b));
  HIPCHK(hipHostMalloc((void**)&b->h_ps, sizeof(PlasticScalars) * ncol));
  HIPCHK(hipHostMalloc((void**)&b->h_bc, sizeof(BCol) * ncol));
  HIPCHK(hipHostMalloc((void**)&b->h_cst, sizeof(int32_t) * ncol));
  HIPCHK(hipHostMalloc((void**)&b->h_stage, sizeof(int32_t) * ncol * STAGE_SLOTS_N));
  b->phase.assign(ncol, BPH_IDLE);
  b->nl_iter.assign(ncol, 0);
  b->cg_tot.assign(ncol, 0);
  b->need_start.assign(ncol, 0);
  b->seeded.assign(ncol, 0);
  b->seed_cg.assign(ncol, 0);
  b->err.assign(ncol, 0);
  b->last_it.assign(ncol, 0);
  b->last_rep.assign(ncol, 0);
  b->cst.assign(ncol, BCOL_OFF);
  b->fmax.assign(ncol, 0.0);
  b->uerr.assign(ncol, 0.0);
  memset(b->h_bc, 0, sizeof(BCol) * ncol);
  b->h_mp.assign((size_t)std::max(1, m->nmat) * ncol, MatParams{});
  HIPCHK(hipMemcpyAsync(b->d_ext, extload, b->n1 * 8, hipMemcpyHostToDevice, s));
  TRY(batch_sync_operator(b));
  HIPCHK(hipMemsetAsync(b->d_mask, 0, ncol * 4, s));
  HIPCHK(hipMemsetAsync(b->d_cst, 0, ncol * 4, s));
  HIPCHK(hipMemsetAsync(b->d_rlist, 0, 65 * 4, s));
  HIPCHK(hipMemsetAsync(b->d_rslot, 0, 64 * 4, s));
  HIPCHK(hipMemsetAsync(b->d_bc, 0, ncol * sizeof(BCol), s));
  HIPCHK(hipMemsetAsync(b->d_ps, 0, ncol * sizeof(PlasticScalars), s));
  HIPCHK(hipMemsetAsync(b->d_x, 0, nv * 8, s));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  *out = guard.release();
  return 0;
}

// steps between x writes (1: every step); a power of 2 dividing the period
static int batch_xlen(const sgeo_batch* b) {
  return b->defer_x ? std::min(b->resid_period, b->pring) : 1;
}

static void batch_drop_graphs(sgeo_batch* b) {
  for (int i = 0; i < 3; i++) {
    if (b->gexec[i]) hipGraphExecDestroy(b->gexec[i]);
    if (b->graph[i]) hipGraphDestroy(b->graph[i]);
    b->gexec[i] = nullptr;
    b->graph[i] = nullptr;
  }
}

extern "C" int sgeo_batch_destroy(sgeo_batch* b) {
  if (!b) return 0;
  hipSetDevice(b->mesh->ctx->device);
  hipStreamSynchronize(b->mesh->ctx->stream);
  void* ptrs[] = {b->d_x,     b->d_oldx,   b->d_load,   b->d_body,  b->d_Y64,
                  b->d_ext,   b->d_xel,    b->d_r,      b->d_p,     b->d_edot,
                  b->d_Y32,   b->d_dinv32, b->d_evpt,   b->d_stress, b->d_BL,
                  b->d_flags, b->d_mp,     b->d_ps,     b->d_bc,    b->d_cst,
                  b->d_mask,  b->d_islast, b->d_cols,   b->d_part,  b->d_vm,
                  b->d_sf,    b->d_nacc,   b->d_rlist,  b->d_rslot, b->d_eorder};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (b->h_ps) hipHostFree(b->h_ps);
  if (b->h_bc) hipHostFree(b->h_bc);
  if (b->h_cst) hipHostFree(b->h_cst);
  if (b->h_stage) hipHostFree(b->h_stage);
  batch_drop_graphs(b);
  delete b;
  return 0;
}

extern "C" int sgeo_batch_set_graphs(sgeo_batch* b, int enable) {
  if (!b) return fail(SGEO_EARG, "null batch");
  b->use_graphs = enable != 0;
  return 0;
}

extern "C" int sgeo_batch_set_residual_period(sgeo_batch* b, int32_t period) {
  if (!b || (period != 1 && period != 2 && period != 4 && period != 8))
    return fail(SGEO_EARG, "sgeo_batch_set_residual_period: period 1, 2, 4 or 8");
  for (int j = 0; j < b->B; j++)
    if (b->phase[j] == BPH_RUN && period != b->resid_period)
      return fail(SGEO_ESTATE, "sgeo_batch_set_residual_period: trials are running");
  if (period != b->resid_period) batch_drop_graphs(b);
  b->resid_period = period;
  return 0;
}

extern "C" int sgeo_batch_set_deferred_x(sgeo_batch* b, int32_t enable) {
  if (!b) return fail(SGEO_EARG, "null batch");
  const bool on = enable != 0;
  for (int j = 0; j < b->B; j++)
    if (b->phase[j] == BPH_RUN && on != b->defer_x)
      return fail(SGEO_ESTATE, "sgeo_batch_set_deferred_x: trials are running");
  if (on != b->defer_x) batch_drop_graphs(b);  // the graphs hold the x-writing steps
  b->defer_x = on;
  return 0;
}

extern "C" int sgeo_batch_set_geometry(sgeo_batch* b, int32_t blocks_per_cu,
                                       int32_t xcd_split) {
  if (!b || blocks_per_cu < 1 || blocks_per_cu > 4 || xcd_split < 0 || xcd_split > 1)
    return fail(SGEO_EARG, "sgeo_batch_set_geometry: blocks_per_cu 1..4, xcd_split 0/1");
  if (blocks_per_cu != b->geom.blocks_per_cu || xcd_split != b->geom.xcd_split)
    batch_drop_graphs(b);
  b->geom.blocks_per_cu = blocks_per_cu;
  b->geom.xcd_split = xcd_split;
  return 0;
}

// Launch-geometry autotuner of the MFMA product: times `reps` launches of
// every candidate (blocks per CU 1..3, XCD-aware split on/off) on the
// engine's own mesh and the columns iterating now, and keeps the fastest.
// The product only rewrites the element products of iterating columns from
// their current search directions, which the next step recomputes before
// any use, so tuning between steps changes no result.
extern "C" int sgeo_batch_autotune(sgeo_batch* b, int32_t reps, int32_t* blocks_per_cu,
                                   int32_t* xcd_split, double* ms_best,
                                   double* ms_table) {
  if (!b || reps < 1) return fail(SGEO_EARG, "sgeo_batch_autotune: bad arguments");
  sgeo_mesh* m = b->mesh;
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int inner = 0;
  for (int j = 0; j < b->B; j++) inner += b->phase[j] == BPH_RUN && b->cst[j] == BCOL_INNER;
  if (!inner) return fail(SGEO_ESTATE, "sgeo_batch_autotune: no column is iterating");
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  double best = 1e30;
  BmvGeom pick = b->geom;
  int k = 0;
  for (int xs = 1; xs >= 0; xs--) {
    for (int bpc = 4; bpc >= 1; bpc--, k++) {
      BmvGeom g;
      g.blocks_per_cu = bpc;
      g.xcd_split = xs;
      auto launch = [&]() {
        launch_bmv_mfma_f32(s, m->nelmt, b->B, c->n_cu, g, m->d_KP32, m->d_gdof,
                            b->d_eorder, b->d_p, b->n1 * b->B, b->d_Y32, b->d_edot,
                            b->d_cst, b->d_bc, BCOL_BIT(BCOL_INNER));
      };
      launch();  // warm-up
      HIPCHK(hipEventRecord(e0, s));
      for (int r = 0; r < reps; r++) launch();
      HIPCHK(hipEventRecord(e1, s));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      const double avg = (double)ms / reps;
      if (ms_table) ms_table[k] = avg;
      if (avg < best) {
        best = avg;
        pick = g;
      }
    }
  }
  HIPCHK(hipGetLastError());
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (pick.blocks_per_cu != b->geom.blocks_per_cu || pick.xcd_split != b->geom.xcd_split)
    batch_drop_graphs(b);  // captured launches carry the old geometry
  b->geom = pick;
  if (blocks_per_cu) *blocks_per_cu = pick.blocks_per_cu;
  if (xcd_split) *xcd_split = pick.xcd_split;
  if (ms_best) *ms_best = best;
  return 0;
}

extern "C" int sgeo_batch_bytes(sgeo_batch* b, size_t* bytes) {
  if (!b || !bytes) return fail(SGEO_EARG, "null argument");
  *bytes = b->bytes;
  return 0;
}

extern "C" int sgeo_batch_begin(sgeo_batch* b, int32_t col, const double* ym,
                                const double* nuf, const double* phif,
                                const double* cohf, const double* psif,
                                double dt, const double* x0, int32_t x0_cg_iter) {
  if (!b || col < 0 || col >= b->B || !ym || !nuf || !phif || !cohf || !psif)
    return fail(SGEO_EARG, "sgeo_batch_begin: bad arguments");
  sgeo_mesh* m = b->mesh;
  sgeo_ctx* c = m->ctx;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  TRY(batch_sync_operator(b));
  const int nmat = std::max(1, m->nmat);
  for (int i = 0; i < m->nmat; i++)
    b->h_mp[(size_t)col * nmat + i] =
        MatParams{ym[i], nuf[i], phif[i], cohf[i], psif[i], dt};
  HIPCHK(hipMemcpyAsync(b->d_mp + (size_t)col * nmat, &b->h_mp[(size_t)col * nmat],
                        nmat * sizeof(MatParams), hipMemcpyHostToDevice, s));
  const size_t ns = (size_t)m->nelmt * 27 * 6;
  // stress_elmt=0, evpt=0, bodyload=0, oldx=0, and x = 0 or the elastic
  // solution (src/slope.F90:249, 281-282)
  HIPCHK(hipMemsetAsync(b->d_evpt + col * ns, 0, ns * 8, s));
  HIPCHK(hipMemsetAsync(b->d_stress + col * ns, 0, ns * 8, s));
  HIPCHK(hipMemsetAsync(b->d_flags + (size_t)col * m->nelmt * 27, 0,
                        (size_t)m->nelmt * 27, s));
  launch_col_io(s, b->n1, b->B, col, b->d_body, nullptr, nullptr);
  launch_col_io(s, b->n1, b->B, col, b->d_oldx, nullptr, nullptr);
  double* d_x0 = nullptr;
  DevScratch tmp;
  int seeded = 0, seed_cg = 0;
  if (x0) {
    TRY(tmp.alloc(&d_x0, (size_t)b->n1));
    HIPCHK(hipMemcpyAsync(d_x0, x0, b->n1 * 8, hipMemcpyHostToDevice, s));
    launch_col_io(s, b->n1, b->B, col, b->d_x, d_x0, nullptr);
    seeded = 1;
    seed_cg = x0_cg_iter;
  } else if (b->memo_valid) {
    launch_col_io(s, b->n1, b->B, col, b->d_x, b->d_xel, nullptr);
    seeded = 1;
    seed_cg = b->memo_cg;
  } else {
    launch_col_io(s, b->n1, b->B, col, b->d_x, nullptr, nullptr);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  b->phase[col] = BPH_RUN;
  b->need_start[col] = 1;
  b->nl_iter[col] = 0;
  b->cg_tot[col] = 0;
  b->err[col] = 0;
  b->fmax[col] = 0.0;
  b->uerr[col] = 0.0;
  b->seeded[col] = seeded;
  b->seed_cg[col] = seed_cg;
  b->cst[col] = BCOL_OFF;
  if (!b->nacc_valid.empty()) b->nacc_valid[col] = 0;
  return 0;
}

// An excavation load increment on column col (src/excavation.F90:509-515):
// materials and dt, evpt, bodyload, x and oldx reset, the stress state set
// from stress0 (host, 6*27*nelmt) or kept from the column's last increment;
// the body then also records each point's plastic strain and stress
// concentration factor at convergence (src/excavation.F90:599-609).  The
// first PCG call solves the increment's load from x = 0, the same system
// for every increment of the stage, so the engine's elastic memo serves it.
extern "C" int sgeo_batch_begin_increment(sgeo_batch* b, int32_t col, const double* ym,
                                          const double* nuf, const double* phif,
                                          const double* cohf, const double* psif,
                                          double dt, const double* stress0) {
  if (!b || col < 0 || col >= b->B)
    return fail(SGEO_EARG, "sgeo_batch_begin_increment: bad arguments");
  sgeo_mesh* m = b->mesh;
  HIPCHK(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  const size_t ns = (size_t)m->nelmt * 27 * 6, np = (size_t)m->nelmt * 27;
  if (!b->d_vm || !b->d_sf) {  // both or neither: the body keys on the pair
    double *vm = b->d_vm, *sf = b->d_sf;
    if (!vm) TRY(dalloc(&vm, np * b->B, &b->bytes));
    b->d_vm = vm;
    if (!sf) TRY(dalloc(&sf, np * b->B, &b->bytes));
    b->d_sf = sf;
  }
  // keep the stress across sgeo_batch_begin's reset
  double* keep = nullptr;
  DevScratch tmp;
  if (!stress0) {
    TRY(tmp.alloc(&keep, ns));
    HIPCHK(hipMemcpyAsync(keep, b->d_stress + col * ns, ns * 8, hipMemcpyDeviceToDevice,
                          s));
  }
  int rc = sgeo_batch_begin(b, col, ym, nuf, phif, cohf, psif, dt, nullptr, 0);
  hipError_t e = hipSuccess;
  if (rc == 0)
    e = stress0 ? hipMemcpyAsync(b->d_stress + col * ns, stress0, ns * 8,
                                 hipMemcpyHostToDevice, s)
                : hipMemcpyAsync(b->d_stress + col * ns, keep, ns * 8,
                                 hipMemcpyDeviceToDevice, s);
  if (rc == 0 && e == hipSuccess) e = hipMemsetAsync(b->d_vm + col * np, 0, np * 8, s);
  if (rc == 0 && e == hipSuccess) e = hipMemsetAsync(b->d_sf + col * np, 0, np * 8, s);
  if (rc == 0 && e == hipSuccess) e = hipStreamSynchronize(s);
  if (rc != 0) return rc;
  if (e != hipSuccess)
    return fail(SGEO_EHIP, "sgeo_batch_begin_increment: %s", hipGetErrorString(e));
  return 0;
}

// The column's stress_elmt (6*27*nelmt, reference layout) to the host.
extern "C" int sgeo_batch_download_stress(sgeo_batch* b, int32_t col, double* stress) {
  if (!b || col < 0 || col >= b->B || !stress) return fail(SGEO_EARG, "bad arguments");
  sgeo_mesh* m = b->mesh;
  HIPCHK(hipSetDevice(m->ctx->device));
  const size_t ns = (size_t)m->nelmt * 27 * 6;
  HIPCHK(hipMemcpyAsync(stress, b->d_stress + col * ns, ns * 8, hipMemcpyDeviceToHost,
                        m->ctx->stream));
  HIPCHK(hipStreamSynchronize(m->ctx->stream));
  return 0;
}

// sgeo_slope_accumulate_nodal for column col of an excavation increment.
extern "C" int sgeo_batch_accumulate_nodal(sgeo_batch* b, int32_t col, int32_t first) {
  if (!b || col < 0 || col >= b->B) return fail(SGEO_EARG, "bad column");
  if (!b->d_vm || !b->d_sf)
    return fail(SGEO_ESTATE, "sgeo_batch_accumulate_nodal: no excavation increment");
  sgeo_mesh* m = b->mesh;
  if (!m->d_nrowptr || !m->d_nslots || m->nnode <= 0)
    return fail(SGEO_ESTATE, "no node map: call sgeo_mesh_set_nodes first");
  HIPCHK(hipSetDevice(m->ctx->device));
  const size_t nn = (size_t)m->nnode, np = (size_t)m->nelmt * 27;
  if (b->d_nacc && b->nacc_nnode != m->nnode) {  // node map replaced since
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    hipFree(b->d_nacc);
    b->bytes -= 2 * (size_t)b->nacc_nnode * b->B * sizeof(double);
    b->d_nacc = nullptr;
  }
  if (!b->d_nacc) {
    b->nacc_valid.assign(b->B, 0);
    b->nacc_init.assign(b->B, 0);
    TRY(dalloc(&b->d_nacc, 2 * nn * b->B, &b->bytes));
    b->nacc_nnode = m->nnode;
  }
  if (!first && !b->nacc_init[col])
    return fail(SGEO_ESTATE,
                "sgeo_batch_accumulate_nodal: first=0 before an accumulation with first=1");
  launch_nodal_accumulate(m->ctx->stream, m->nnode, m->d_nrowptr, m->d_nslots,
                          b->d_vm + col * np, b->d_sf + col * np, b->d_nacc + col * 2 * nn,
                          first ? 1 : 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(m->ctx->stream));
  b->nacc_valid[col] = 1;
  b->nacc_init[col] = 1;
  return 0;
}

// A released column stops at once: its device state is set to OFF too
// (queued behind the steps already launched), so a column abandoned in the
// middle of a PCG call costs no further vector passes or FP64 residuals.
extern "C" int sgeo_batch_release(sgeo_batch* b, int32_t col) {
  if (!b || col < 0 || col >= b->B) return fail(SGEO_EARG, "bad column");
  b->phase[col] = BPH_IDLE;
  b->need_start[col] = 0;
  b->cst[col] = BCOL_OFF;
  b->h_cst[col] = BCOL_OFF;
  HIPCHK(hipSetDevice(b->mesh->ctx->device));
  launch_set_i32(b->mesh->ctx->stream, b->d_cst + col, BCOL_OFF);
  HIPCHK(hipGetLastError());
  return 0;
}

// Small per-call uploads (column masks, lists, states) go through pinned
// staging slots and are queued on the stream without a synchronisation; a
// slot is rewritten only after a synchronisation has retired its previous
// copy (stage_busy marks slots whose copy may still be queued).
enum StageSlot { STAGE_START_MASK = 0, STAGE_CST, STAGE_BODY_MASK, STAGE_ISLAST,
                 STAGE_COLS, STAGE_SLOTS };

static_assert(STAGE_SLOTS == STAGE_SLOTS_N, "staging slots");

static int stage_upload(sgeo_batch* b, int slot, void* dst, const int32_t* src, int n) {
  hipStream_t s = b->mesh->ctx->stream;
  if (b->stage_busy & (1u << slot)) {
    HIPCHK(hipStreamSynchronize(s));
    b->stage_busy = 0;
  }
  int32_t* h = b->h_stage + (size_t)slot * b->B;
  memcpy(h, src, (size_t)n * 4);
  HIPCHK(hipMemcpyAsync(dst, h, (size_t)n * 4, hipMemcpyHostToDevice, s));
  b->stage_busy |= 1u << slot;
  return 0;
}

static int upload_mask(sgeo_batch* b, int slot, const std::vector<int32_t>& cols) {
  std::vector<int32_t> mask(b->B, 0);
  for (int j : cols) mask[j] = 1;
  return stage_upload(b, slot, b->d_mask, mask.data(), b->B);
}

// Starts a PCG call for `cols`: load = extload + bodyload, fresh scalars;
// a column seeded with the elastic solution skips its first call.
static int batch_start(sgeo_batch* b, const std::vector<int32_t>& cols, double cg_tol,
                       int32_t cg_maxiter, double inner_tol) {
  hipStream_t s = b->mesh->ctx->stream;
  TRY(upload_mask(b, STAGE_START_MASK, cols));
  launch_bload(s, b->n1, b->B, b->d_ext, b->d_body, b->d_load, b->d_mask);
  for (int j : cols) {
    const bool skip = b->seeded[j] && b->nl_iter[j] == 0;
    b->cst[j] = skip ? BCOL_DONE : BCOL_START;
    BCol& bc = b->h_bc[j];
    bc = BCol{};
    bc.tol = cg_tol;
    bc.inner_tol = inner_tol;
    bc.maxiter = cg_maxiter;
    bc.xlen = batch_xlen(b);
    bc.pmask = b->pring - 1;
    bc.it = skip ? b->seed_cg[j] : 0;
    b->need_start[j] = 0;
    b->last_it[j] = bc.it;
    b->last_rep[j] = 0;
  }
  TRY(stage_upload(b, STAGE_CST, b->d_cst, b->cst.data(), b->B));
  HIPCHK(hipMemcpyAsync(b->d_bc, b->h_bc, b->B * sizeof(BCol), hipMemcpyHostToDevice, s));
  return 0;
}

static bool pcg_running(int st) {
  return st == BCOL_START || st == BCOL_INNER || st == BCOL_REPLACE;
}

// One CG step of every column (launch sequence; no memset, no host work).
// With `resid` (a residual step), the step first forms the FP64 residual of
// the columns that start a call or are due for a replacement (a no-op
// launch when there are none).
static void batch_step_launches(sgeo_batch* b, bool prof, bool resid, bool xw) {
  sgeo_mesh* m = b->mesh;
  sgeo_ctx* c = m->ctx;
  hipStream_t s = c->stream;
  const int B = b->B;
  const int64_t n1 = b->n1;
  const int64_t nch = batch_equation_chunks(n1);
  if (resid) {
    launch_bpromote(s, B, b->d_cst, b->d_bc, b->d_rlist, b->d_rslot);
    launch_bmv_f64(s, m->nelmt, B, c->n_cu, m->d_KP64, m->d_gdof, b->d_x, b->d_Y64,
                   b->d_rlist);
    launch_bresid(s, n1, B, m->d_rowptr, m->d_slots, b->d_Y64, b->d_load, b->d_dinv32,
                  b->d_x, b->d_r, b->d_p, b->d_part, b->d_cst, b->d_rslot);
    launch_bfinalize(s, nch, B, b->d_part, b->d_cst, b->d_bc, BFIN_RESID);
  }
  launch_bpupdate(s, n1, B, b->d_r, b->d_dinv32, b->d_p, b->d_part, b->d_cst, b->d_bc);
  if (prof) prof_mark(c);
  // Y = K P and the elements' p.Kp shares; alpha = r.z / p.Kp
  launch_bmv_mfma_f32(s, m->nelmt, B, c->n_cu, b->geom, m->d_KP32, m->d_gdof, b->d_eorder,
                      b->d_p, n1 * B, b->d_Y32, b->d_edot, b->d_cst, b->d_bc,
                      BCOL_BIT(BCOL_INNER));
  if (prof) prof_mark(c);
  launch_bfinalize(s, m->nelmt, B, b->d_edot, b->d_cst, b->d_bc, BFIN_ALPHA);
  // kp assembled from Y; r and r.z; on the x-writing steps x and max|x|;
  // the stop test
  launch_bupdate(s, n1, B, m->d_rowptr, m->d_slots, b->d_Y32, b->d_dinv32, b->d_r,
                 b->d_part, b->d_cst, b->d_bc);
  if (xw) launch_bxfold(s, n1, B, b->d_x, b->d_p, b->d_part, b->d_cst, b->d_bc);
  launch_bfinalize(s, nch, B, b->d_part, b->d_cst, b->d_bc, BFIN_INNER, xw ? 1 : 0);
}

// `chunk` lockstep CG steps of every running column (FP64 residuals on the
// engine's residual steps), then reads the states back; h_bc holds the
// device scalars on return.  Graphs (8, 16 or 32 steps, multiples of the
// period) are replayed only from a step aligned to the period, so their
// residual positions are the engine's.
static int batch_steps(sgeo_batch* b, int chunk, bool* any_running) {
  sgeo_mesh* m = b->mesh;
  sgeo_ctx* c = m->ctx;
  hipStream_t s = c->stream;
  const int B = b->B;
  const auto t0 = std::chrono::steady_clock::now();
  const int gi = chunk == 8 ? 0 : chunk == 16 ? 1 : chunk == 32 ? 2 : -1;
  const int per = b->resid_period;
  const int xlen = batch_xlen(b);  // divides per
  if (b->use_graphs && !c->prof && gi >= 0 && b->gstep % per == 0) {
    if (!b->gexec[gi]) {
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      for (int it = 0; it < chunk; it++)
        batch_step_launches(b, false, it % per == 0, (it + 1) % xlen == 0);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(s, &g);
      if (ec != hipSuccess || !g) return fail(SGEO_EHIP, "graph capture failed");
      b->graph[gi] = g;
      HIPCHK(hipGraphInstantiate(&b->gexec[gi], g, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(b->gexec[gi], s));
  } else {
    for (int it = 0; it < chunk; it++)
      batch_step_launches(b, c->prof, (b->gstep + it) % per == 0,
                          (b->gstep + it + 1) % xlen == 0);
  }
  b->gstep += chunk;
  b->stat_steps += chunk;
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->h_cst, b->d_cst, B * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(b->h_bc, b->d_bc, B * sizeof(BCol), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  b->stage_busy = 0;
  {
    const double us = std::chrono::duration<double, std::micro>(
                          std::chrono::steady_clock::now() - t0).count() / chunk;
    b->step_us = b->step_us > 0.0 ? 0.8 * b->step_us + 0.2 * us : us;
  }
  if (c->prof) TRY(prof_flush(c));
  bool running = false;
  for (int j = 0; j < B; j++) {
    if (b->phase[j] != BPH_RUN) continue;
    if (b->cst[j] == BCOL_OFF) continue;  // waiting for its start
    b->stat_inner += b->h_bc[j].it - b->last_it[j];
    b->stat_outer += b->h_bc[j].nrep - b->last_rep[j];
    b->last_it[j] = b->h_bc[j].it;
    b->last_rep[j] = b->h_bc[j].nrep;
    b->cst[j] = b->h_cst[j];
    if (b->h_bc[j].xerr)
      return fail(SGEO_ESTATE, "batch column %d: x update out of phase with its iterations",
                  j);
    running = running || pcg_running(b->cst[j]);
  }
  *any_running = running;
  return 0;
}

// PCG finished for `cols` (DONE or FAIL): keeps the elastic memo, then
// x(0)=0, uerr and nl_isconv, the element x GLL body and the body load
// (src/slope.F90:297-391) for those columns.  Updates nl_iter, cg_tot,
// fmax, uerr and err; the columns are left with cst OFF.
static int batch_body(sgeo_batch* b, const std::vector<int32_t>& cols, double nl_tol,
                      int32_t nl_maxiter, const int32_t* is_last_override) {
  sgeo_mesh* m = b->mesh;
  hipStream_t s = m->ctx->stream;
  const int B = b->B;
  // x of the solves that ended between two x writes
  launch_bmaterialize(s, b->n1, B, b->d_x, b->d_p, b->d_cst, b->d_bc);
  for (int j : cols) {  // first PCG call of a trial from x = 0: the elastic memo
    if (!b->memo_valid && b->nl_iter[j] == 0 && !b->seeded[j] && b->cst[j] == BCOL_DONE) {
      launch_col_io(s, b->n1, B, j, b->d_x, nullptr, b->d_xel);
      b->memo_valid = true;
      b->memo_cg = b->h_bc[j].it;
    }
  }
  std::vector<int32_t> islast(B, 0);
  for (int j : cols)
    islast[j] = is_last_override ? (is_last_override[j] ? 1 : 0)
                                 : (b->nl_iter[j] + 1 == nl_maxiter ? 1 : 0);
  TRY(upload_mask(b, STAGE_BODY_MASK, cols));
  TRY(stage_upload(b, STAGE_ISLAST, b->d_islast, islast.data(), B));
  TRY(stage_upload(b, STAGE_COLS, b->d_cols, cols.data(), (int)cols.size()));
  launch_buerr(s, b->n1, B, b->d_x, b->d_oldx, b->d_part, b->d_ps, b->d_mask, nl_tol);
  BodyArgs ba;
  ba.nel = m->nelmt;
  ba.corners = m->d_corners;
  ba.gdof = m->d_gdof;
  ba.mat_id = m->d_mat;
  ba.mp = b->d_mp;
  ba.nmat = std::max(1, m->nmat);
  ba.x = b->d_x;
  ba.xs = B;
  ba.evpt = b->d_evpt;
  ba.stress = b->d_stress;
  ba.BL = b->d_BL;
  ba.flags = b->d_flags;
  ba.ps = b->d_ps;
  ba.is_last_col = b->d_islast;
  ba.cols = b->d_cols;
  ba.ncols = (int)cols.size();
  ba.excav = b->d_vm && b->d_sf;  // per-point vmeps and scf at convergence
  ba.vm_pt = b->d_vm;
  ba.sf_pt = b->d_sf;
  launch_plastic_body(s, ba);
  launch_bbody_assemble(s, b->n1, B, m->nelmt, m->d_rowptr, m->d_slots, b->d_BL,
                        b->d_body, b->d_mask, b->d_islast, b->d_ps);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->h_ps, b->d_ps, B * sizeof(PlasticScalars),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  b->stage_busy = 0;
  for (int j : cols) {
    b->nl_iter[j] += 1;
    b->cg_tot[j] += b->h_bc[j].it;
    if (b->cst[j] != BCOL_DONE) b->err[j] = -1;
    b->uerr[j] = b->h_ps[j].uerr;
    unsigned long long bits = b->h_ps[j].fmax_bits;
    memcpy(&b->fmax[j], &bits, 8);
    b->cst[j] = BCOL_OFF;
  }
  return 0;
}

// One plastic iteration of every running column in lockstep
// (src/slope.F90:286-391).
extern "C" int sgeo_batch_plastic_iter(sgeo_batch* b, const int32_t* is_last,
                                       double cg_tol, int32_t cg_maxiter,
                                       double inner_tol, double nl_tol,
                                       int32_t* cg_iter, int32_t* errcode,
                                       double* uerr, int32_t* nl_isconv,
                                       double* fmax) {
  if (!b || !is_last) return fail(SGEO_EARG, "sgeo_batch_plastic_iter: bad arguments");
  sgeo_ctx* c = b->mesh->ctx;
  HIPCHK(hipSetDevice(c->device));
  const int B = b->B;
  std::vector<int32_t> cols;
  for (int j = 0; j < B; j++)
    if (b->phase[j] == BPH_RUN) cols.push_back(j);
  if (cols.empty()) return fail(SGEO_ESTATE, "no active column");
  TRY(batch_sync_operator(b));
  std::vector<int32_t> cg0(B, 0);
  for (int j : cols) cg0[j] = b->cg_tot[j];
  TRY(batch_start(b, cols, cg_tol, cg_maxiter, inner_tol));
  const int64_t max_steps = 2 * (int64_t)cg_maxiter + 64;
  int64_t steps = 0;
  bool running = false;
  for (int j : cols) running = running || pcg_running(b->cst[j]);
  for (int chunk = 8; running; chunk = std::min(chunk * 2, 32)) {
    TRY(batch_steps(b, chunk, &running));
    steps += chunk;
    if (running && steps > max_steps)  // every step advances each running column
      return fail(SGEO_ESTATE, "batch PCG did not terminate in %lld steps",
                  (long long)steps);
  }
  TRY(batch_body(b, cols, nl_tol, 0, is_last));
  for (int j : cols) {
    if (cg_iter) cg_iter[j] = b->cg_tot[j] - cg0[j];
    if (errcode) errcode[j] = b->err[j];
    if (uerr) uerr[j] = b->uerr[j];
    if (nl_isconv) nl_isconv[j] = b->h_ps[j].nl_isconv;
    if (fmax) fmax[j] = b->fmax[j];
    b->err[j] = 0;
  }
  return 0;
}

// Runs every running column through its own sequence of plastic iterations
// until at least one trial ends (nl_isconv, nl_maxiter reached, or a PCG
// call that did not converge), the elastic memo becomes available, or
// max_steps CG steps have run.
extern "C" int sgeo_batch_advance(sgeo_batch* b, double cg_tol, int32_t cg_maxiter,
                                  double inner_tol, double nl_tol, int32_t nl_maxiter,
                                  int64_t max_steps, int32_t* ended, int32_t* nl_tot,
                                  int32_t* cg_tot, double* fmax, int32_t* errcode,
                                  int32_t* memo_ready) {
  if (!b || !ended || nl_maxiter < 1 || max_steps < 1)
    return fail(SGEO_EARG, "sgeo_batch_advance: bad arguments");
  sgeo_ctx* c = b->mesh->ctx;
  HIPCHK(hipSetDevice(c->device));
  TRY(batch_sync_operator(b));
  const int B = b->B;
  for (int j = 0; j < B; j++) ended[j] = 0;
  const bool memo0 = b->memo_valid;
  int64_t steps = 0;
  int chunk = (b->step_us > 0.0 && b->step_us < 1000.0) ? 32 : 8;
  for (;;) {
    // columns that need a new PCG call
    std::vector<int32_t> starts;
    for (int j = 0; j < B; j++)
      if (b->phase[j] == BPH_RUN && b->need_start[j]) starts.push_back(j);
    if (!starts.empty()) {
      TRY(batch_start(b, starts, cg_tol, cg_maxiter, inner_tol));
      // Long steps: short first chunks, so a column that ends its call
      // soon after a start waits little.  Short steps (small meshes, where
      // each synchronisation and body pass costs more than a step): whole
      // 32-step chunks, so ends are handled together.
      chunk = (b->step_us > 0.0 && b->step_us < 1000.0) ? 32 : 8;
    }
    bool running = false;
    for (int j = 0; j < B; j++)
      running = running || (b->phase[j] == BPH_RUN && pcg_running(b->cst[j]));
    if (running) {
      // never past max_steps: whole graphs of 32, 16 or 8 steps while they
      // fit, the last few steps launched one by one
      int64_t run = std::min<int64_t>(chunk, max_steps - steps);
      run = run >= 32 ? 32 : run >= 16 ? 16 : run >= 8 ? 8 : run;
      TRY(batch_steps(b, (int)run, &running));
      steps += run;
      chunk = std::min(chunk * 2, 32);
    }
    // columns whose PCG call finished: body pass
    std::vector<int32_t> fin;
    for (int j = 0; j < B; j++)
      if (b->phase[j] == BPH_RUN && (b->cst[j] == BCOL_DONE || b->cst[j] == BCOL_FAIL))
        fin.push_back(j);
    bool any_end = false;
    if (!fin.empty()) {
      TRY(batch_body(b, fin, nl_tol, nl_maxiter, nullptr));
      for (int j : fin) {
        const bool end = b->err[j] != 0 || b->h_ps[j].nl_isconv ||
                         b->nl_iter[j] >= nl_maxiter;
        if (end) {
          b->phase[j] = BPH_ENDED;
          ended[j] = 1;
          any_end = true;
          if (nl_tot) nl_tot[j] = b->nl_iter[j];
          if (cg_tot) cg_tot[j] = b->cg_tot[j];
          if (fmax) fmax[j] = b->fmax[j];
          if (errcode) errcode[j] = b->err[j];
        } else {
          b->need_start[j] = 1;
        }
      }
    }
    bool busy = false;
    for (int j = 0; j < B; j++) busy = busy || b->phase[j] == BPH_RUN;
    if (any_end || !busy || (b->memo_valid && !memo0) || steps >= max_steps) break;
  }
  if (memo_ready) *memo_ready = b->memo_valid ? 1 : 0;
  return 0;
}

// Plastic iterations each column's current trial has done (host copy,
// as of the last body pass; 0 for columns without a running trial).
extern "C" int sgeo_batch_nl_iters(sgeo_batch* b, int32_t* nl_iter) {
  if (!b || !nl_iter) return fail(SGEO_EARG, "sgeo_batch_nl_iters: null argument");
  for (int j = 0; j < b->B; j++) nl_iter[j] = b->phase[j] == BPH_RUN ? b->nl_iter[j] : 0;
  return 0;
}

extern "C" int sgeo_batch_download(sgeo_batch* b, int32_t col, double* x,
                                   uint8_t* flags) {
  if (!b || col < 0 || col >= b->B) return fail(SGEO_EARG, "bad column");
  sgeo_mesh* m = b->mesh;
  hipStream_t s = m->ctx->stream;
  HIPCHK(hipSetDevice(m->ctx->device));
  if (x) {
    double* tmp = nullptr;
    TRY(dalloc(&tmp, (size_t)b->n1));
    launch_col_io(s, b->n1, b->B, col, b->d_x, nullptr, tmp);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(x, tmp, b->n1 * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(tmp);
    if (e != hipSuccess)
      return fail(SGEO_EHIP, "sgeo_batch_download: %s", hipGetErrorString(e));
  }
  if (flags)
    HIPCHK(hipMemcpyAsync(flags, b->d_flags + (size_t)col * m->nelmt * 27,
                          (size_t)m->nelmt * 27, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int sgeo_batch_nodal_fields(sgeo_batch* b, int32_t col, float* stress,
                                       float* psigma, float* nsigma, float* scf,
                                       float* vmeps) {
  if (!b || col < 0 || col >= b->B) return fail(SGEO_EARG, "bad column");
  const size_t ns = (size_t)b->mesh->nelmt * 27 * 6;
  const double* acc = (b->d_nacc && b->nacc_nnode == b->mesh->nnode && b->nacc_valid[col])
                          ? b->d_nacc + (size_t)col * 2 * b->mesh->nnode
                          : nullptr;
  return nodal_fields(b->mesh, b->d_stress + (size_t)col * ns, acc, stress, psigma,
                      nsigma, scf, vmeps);
}

extern "C" int sgeo_batch_stats(sgeo_batch* b, int64_t* inner_iters,
                                int64_t* outer_iters, int64_t* steps) {
  if (!b) return fail(SGEO_EARG, "null batch");
  if (inner_iters) *inner_iters = b->stat_inner;
  if (outer_iters) *outer_iters = b->stat_outer;
  if (steps) *steps = b->stat_steps;
  return 0;
}

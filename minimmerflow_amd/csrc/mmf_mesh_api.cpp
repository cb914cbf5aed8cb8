// mmf_mesh_api.cpp -- C-ABI of the explicit-interface-list path (include/mmf.h,
// mmf_mesh_*): the general MeshGeometricalInfo (mesh_info.hpp:31-105) as a
// face list, the per-face BC storage of computation() (main.cpp:251-277), the
// RHS (euler.cpp:127-249) and the RK3 loop (main.cpp:377-506).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "mmf.h"
#include "mmf_internal.hpp"
#include "mmf_mesh.hpp"

using mmf::Ctrl;
using mmf::MeshDev;

struct mmf_mesh {
  MeshDev m{};
  int32_t dim = 3;
  double hmin = 0.0;
  hipStream_t stream = nullptr;
  double *dU = nullptr, *dW = nullptr, *dR = nullptr, *dAF = nullptr;
  int32_t *downer = nullptr, *dneigh = nullptr, *dkind = nullptr, *doff = nullptr, *dinc = nullptr;
  double *darea = nullptr, *dnormal = nullptr, *dinfo = nullptr, *dvol = nullptr, *darea_tab = nullptr;
  uint8_t *dflags = nullptr;
  Ctrl *dctrl = nullptr;
  Ctrl *hctrl = nullptr;
  // host copies needed to rebuild the face kinds when the cell flags change:
  // border faces' kinds, every face's geometry bits (normal code, area slot)
  std::vector<int32_t> owner, neigh, border_kind, geo;
  std::string err;
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> fev, cev;
  double face_ms = 0.0, cell_ms = 0.0;
  int64_t stages = 0;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  bool use_graphs = true;
};

namespace {

int mfail(mmf_mesh *m, int code, const std::string &msg) {
  if (m) m->err = msg;
  return code;
}

#define MCHK(m, expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return mfail(m, MMF_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

void drop_graphs(mmf_mesh *m) {
  for (int i = 0; i < 2; ++i) {
    if (m->exec[i]) (void)hipGraphExecDestroy(m->exec[i]);
    if (m->graph[i]) (void)hipGraphDestroy(m->graph[i]);
    m->exec[i] = nullptr;
    m->graph[i] = nullptr;
  }
}

int read_ctrl(mmf_mesh *m) {
  MCHK(m, hipMemcpyAsync(m->hctrl, m->dctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return MMF_OK;
}

int nonfinite_status(mmf_mesh *m) {
  if (!m->hctrl->nonfinite) return MMF_OK;
  char buf[160];
  std::snprintf(buf, sizeof(buf), "%llu step(s) met a non-finite state (NaN or infinity); t stays at %.17g",
                (unsigned long long)m->hctrl->nonfinite, m->hctrl->t);
  return mfail(m, MMF_ERR_STATE, buf);
}

// Face kinds from the border kinds and the fluid flags: an interior face is a
// wall where exactly one side is fluid (main.cpp:271-275).
int upload_kinds(mmf_mesh *m, const std::vector<uint8_t> &fluid) {
  const int64_t F = m->m.F;
  std::vector<int32_t> k((size_t)F);
  for (int64_t f = 0; f < F; ++f) {
    const int32_t nb = m->neigh[(size_t)f];
    if (nb < 0) {
      k[(size_t)f] = m->border_kind[(size_t)f] | m->geo[(size_t)f];
    } else {
      const bool wall = fluid.empty() ? false : (fluid[(size_t)m->owner[(size_t)f]] != fluid[(size_t)nb]);
      k[(size_t)f] = (wall ? (MMF_BC_WALL + 1) : 0) | m->geo[(size_t)f];
    }
  }
  MCHK(m, hipMemcpy(m->dkind, k.data(), sizeof(int32_t) * (size_t)F, hipMemcpyHostToDevice));
  return MMF_OK;
}

int launch_stage(mmf_mesh *m, int stage, int mode) {
  const int slot = (mode == mmf::MODE_RHS_ONLY) ? 3 : stage - 1;
  const double *X = (stage == 1) ? m->dU : m->dW;
  hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
  if (m->timing)
    for (auto &x : e) MCHK(m, hipEventCreate(&x));
  if (m->timing) MCHK(m, hipEventRecord(e[0], m->stream));
  hipError_t r = mmf::launch_mesh_faces(m->m, X, m->dAF, m->dctrl, slot, mode, m->stream);
  if (r != hipSuccess) return mfail(m, MMF_ERR_HIP, std::string("face kernel: ") + hipGetErrorString(r));
  if (m->timing) {
    MCHK(m, hipEventRecord(e[1], m->stream));
    MCHK(m, hipEventRecord(e[2], m->stream));
  }
  double *out = (mode == mmf::MODE_RHS_ONLY) ? m->dR : (stage == 3 ? m->dU : m->dW);
  r = mmf::launch_mesh_cells(m->m, stage, m->dU, m->dW, out, m->dAF, m->dctrl, mode, m->stream);
  if (r != hipSuccess) return mfail(m, MMF_ERR_HIP, std::string("cell kernel: ") + hipGetErrorString(r));
  if (m->timing) {
    MCHK(m, hipEventRecord(e[3], m->stream));
    m->fev.emplace_back(e[0], e[1]);
    m->cev.emplace_back(e[2], e[3]);
  }
  return MMF_OK;
}

int enqueue_steps(mmf_mesh *m, int64_t n) {
  for (int64_t s = 0; s < n; ++s) {
    for (int stage = 1; stage <= 3; ++stage) {
      int rc = launch_stage(m, stage, mmf::MODE_STEP);
      if (rc != MMF_OK) return rc;
    }
    hipError_t e = mmf::launch_advance(m->dctrl, m->stream);
    if (e != hipSuccess) return mfail(m, MMF_ERR_HIP, std::string("advance: ") + hipGetErrorString(e));
  }
  return MMF_OK;
}

constexpr int GRAPH_STEPS = 16;

int capture(mmf_mesh *m, int slot, int64_t n) {
  MCHK(m, hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_steps(m, n);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(m->stream, &g);
  if (rc != MMF_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return mfail(m, MMF_ERR_HIP, std::string("end of step capture: ") + hipGetErrorString(e));
  m->graph[slot] = g;
  MCHK(m, hipGraphInstantiate(&m->exec[slot], g, nullptr, nullptr, 0));
  return MMF_OK;
}

int collect_timing(mmf_mesh *m) {
  for (int which = 0; which < 2; ++which) {
    auto &v = which ? m->cev : m->fev;
    for (auto &p : v) {
      MCHK(m, hipEventSynchronize(p.second));
      float ms = 0.f;
      MCHK(m, hipEventElapsedTime(&ms, p.first, p.second));
      (which ? m->cell_ms : m->face_ms) += ms;
      if (which) m->stages += 1;
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
    v.clear();
  }
  return MMF_OK;
}

template <typename T>
int upload(mmf_mesh *m, T **dst, const T *src, size_t n) {
  MCHK(m, hipMalloc(dst, sizeof(T) * (n ? n : 1)));
  if (n) MCHK(m, hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
  return MMF_OK;
}

}  // namespace

extern "C" {

const char *mmf_mesh_last_error(const mmf_mesh *m) { return m ? m->err.c_str() : "null mesh"; }

int mmf_mesh_destroy(mmf_mesh *m) {
  if (!m) return MMF_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (auto *v : {&m->fev, &m->cev})
    for (auto &p : *v) {
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
  drop_graphs(m);  // before the buffers their kernel nodes point at
  for (void *p : {(void *)m->dU, (void *)m->dW, (void *)m->dR, (void *)m->dAF, (void *)m->downer, (void *)m->dneigh,
                  (void *)m->dkind, (void *)m->doff, (void *)m->dinc, (void *)m->darea, (void *)m->dnormal,
                  (void *)m->dinfo, (void *)m->dvol, (void *)m->dflags, (void *)m->dctrl, (void *)m->darea_tab})
    if (p) (void)hipFree(p);
  if (m->hctrl) (void)hipHostFree(m->hctrl);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return MMF_OK;
}

int mmf_mesh_create(const mmf_mesh_desc *d, mmf_mesh **out) {
  if (!d || !out) return MMF_ERR_INVALID;
  *out = nullptr;
  if (d->dim != 2 && d->dim != 3) return MMF_ERR_INVALID;
  if (d->ncells < 1 || d->ncells >= ((int64_t)1 << 31) || d->nfaces < 1 || d->nfaces >= ((int64_t)1 << 30))
    return MMF_ERR_INVALID;
  if (!d->cell_volume || !d->face_owner || !d->face_neigh || !d->face_area || !d->face_normal) return MMF_ERR_INVALID;
  if (!(d->min_cell_size > 0.0)) return MMF_ERR_INVALID;
  const int64_t N = d->ncells, F = d->nfaces;
  for (int64_t c = 0; c < N; ++c)
    if (!(d->cell_volume[c] > 0.0)) return MMF_ERR_INVALID;
  // Dirichlet data slots (distinct states)
  std::vector<double> info;
  std::map<std::vector<double>, int32_t> slots;
  std::vector<int32_t> owner((size_t)F), neigh((size_t)F), bkind((size_t)F, 0);
  std::vector<int32_t> count((size_t)N + 1, 0);
  for (int64_t f = 0; f < F; ++f) {
    const int64_t o = d->face_owner[f], nb = d->face_neigh[f];
    if (o < 0 || o >= N || nb < -1 || nb >= N || nb == o) return MMF_ERR_INVALID;
    if (!(d->face_area[f] >= 0.0)) return MMF_ERR_INVALID;
    owner[(size_t)f] = (int32_t)o;
    neigh[(size_t)f] = (int32_t)nb;
    count[(size_t)o + 1] += 1;
    if (nb >= 0) {
      count[(size_t)nb + 1] += 1;
      continue;
    }
    const int32_t bc = d->face_bc ? d->face_bc[f] : MMF_BC_FREE_FLOW;
    if (bc < MMF_BC_FREE_FLOW || bc > MMF_BC_DIRICHLET) return MMF_ERR_INVALID;
    int32_t slot = 0;
    if (bc == MMF_BC_DIRICHLET) {
      if (!d->face_info) return MMF_ERR_INVALID;
      std::vector<double> key(d->face_info + 5 * f, d->face_info + 5 * f + 5);
      auto it = slots.find(key);
      if (it == slots.end()) {
        slot = (int32_t)slots.size();
        slots.emplace(key, slot);
        info.insert(info.end(), key.begin(), key.end());
      } else {
        slot = it->second;
      }
    }
    bkind[(size_t)f] = (bc + 1) | (slot << mmf::MESH_KIND_SLOT_SHIFT);
  }
  if (slots.size() > (size_t)mmf::MESH_MAX_SLOTS) return MMF_ERR_INVALID;
  // geometry bits: axis-aligned unit normals (exactly +-1 and +0) as a code,
  // areas from a table of up to 255 distinct values; the rest from arrays
  std::vector<int32_t> geo((size_t)F);
  std::vector<double> area_tab;
  std::map<uint64_t, int32_t> area_slot;
  bool need_area = false, need_normal = false;
  for (int64_t f = 0; f < F; ++f) {
    const double *nv = d->face_normal + 3 * f;
    int32_t code = mmf::MESH_NRM_GENERAL;
    for (int a = 0; a < 3 && code == mmf::MESH_NRM_GENERAL; ++a) {
      bool others_zero = true;
      for (int b = 0; b < 3; ++b) {
        uint64_t bits;
        std::memcpy(&bits, &nv[b], sizeof(bits));
        if (b != a && bits != 0) others_zero = false;  // +0.0 only
      }
      if (others_zero && nv[a] == 1.0) code = 2 * a;
      else if (others_zero && nv[a] == -1.0) code = 2 * a + 1;
    }
    need_normal = need_normal || code == mmf::MESH_NRM_GENERAL;
    uint64_t abits;
    std::memcpy(&abits, &d->face_area[f], sizeof(abits));
    int32_t as = mmf::MESH_AREA_GENERAL;
    auto it = area_slot.find(abits);
    if (it != area_slot.end()) {
      as = it->second;
    } else if ((int)area_tab.size() < mmf::MESH_MAX_AREAS) {
      as = (int32_t)area_tab.size();
      area_slot.emplace(abits, as);
      area_tab.push_back(d->face_area[f]);
    }
    need_area = need_area || as == mmf::MESH_AREA_GENERAL;
    geo[(size_t)f] = (code << mmf::MESH_KIND_NRM_SHIFT) | (as << mmf::MESH_KIND_AREA_SHIFT);
  }
  // incidences per cell in increasing face index: the order the interface
  // loop (euler.cpp:153-247) adds them to each residual
  for (int64_t c = 0; c < N; ++c) count[(size_t)c + 1] += count[(size_t)c];
  std::vector<int32_t> inc((size_t)count[(size_t)N]);
  {
    std::vector<int32_t> pos(count.begin(), count.end() - 1);
    for (int64_t f = 0; f < F; ++f) {
      inc[(size_t)pos[(size_t)owner[(size_t)f]]++] = (int32_t)(f << 1);
      if (neigh[(size_t)f] >= 0) inc[(size_t)pos[(size_t)neigh[(size_t)f]]++] = (int32_t)((f << 1) | 1);
    }
  }
  std::vector<double> nrm;
  if (need_normal) {
    nrm.resize((size_t)(3 * F));
    for (int64_t f = 0; f < F; ++f)
      for (int a = 0; a < 3; ++a) nrm[(size_t)(a * F + f)] = d->face_normal[3 * f + a];
  }

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MMF_ERR_NO_DEVICE;
  mmf_mesh *m = new (std::nothrow) mmf_mesh();
  if (!m) return MMF_ERR_OUT_OF_MEMORY;
  auto cleanup = [&](int code) {
    mmf_mesh_destroy(m);
    return code;
  };
  m->dim = d->dim;
  m->hmin = d->min_cell_size;
  m->m.N = N;
  m->m.F = F;
  m->m.fs = (N + 63) / 64 * 64;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(MMF_ERR_HIP);
  const size_t sb = sizeof(double) * 5 * (size_t)m->m.fs;
  if (hipMalloc(&m->dU, sb) != hipSuccess || hipMalloc(&m->dW, sb) != hipSuccess ||
      hipMalloc(&m->dR, sb) != hipSuccess || hipMalloc(&m->dAF, sizeof(double) * 5 * (size_t)F) != hipSuccess ||
      hipMalloc(&m->dkind, sizeof(int32_t) * (size_t)F) != hipSuccess || hipMalloc(&m->dctrl, sizeof(Ctrl)) != hipSuccess ||
      hipHostMalloc(&m->hctrl, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess)
    return cleanup(MMF_ERR_OUT_OF_MEMORY);
  if (upload(m, &m->downer, owner.data(), (size_t)F) || upload(m, &m->dneigh, neigh.data(), (size_t)F) ||
      (need_area && upload(m, &m->darea, d->face_area, (size_t)F)) ||
      (need_normal && upload(m, &m->dnormal, nrm.data(), nrm.size())) ||
      upload(m, &m->darea_tab, area_tab.data(), area_tab.size()) ||
      upload(m, &m->dinfo, info.data(), info.size()) || upload(m, &m->dvol, d->cell_volume, (size_t)N) ||
      upload(m, &m->doff, count.data(), (size_t)N + 1) || upload(m, &m->dinc, inc.data(), inc.size()))
    return cleanup(MMF_ERR_OUT_OF_MEMORY);
  m->owner.swap(owner);
  m->neigh.swap(neigh);
  m->border_kind.swap(bkind);
  m->geo.swap(geo);
  if (upload_kinds(m, {}) != MMF_OK) return cleanup(MMF_ERR_HIP);
  if (hipMemset(m->dU, 0, sb) != hipSuccess || hipMemset(m->dW, 0, sb) != hipSuccess ||
      hipMemset(m->dR, 0, sb) != hipSuccess || hipMemset(m->dctrl, 0, sizeof(Ctrl)) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess)
    return cleanup(MMF_ERR_HIP);
  std::memset(m->hctrl, 0, sizeof(Ctrl));
  m->m.owner = m->downer;
  m->m.neigh = m->dneigh;
  m->m.area = m->darea;
  m->m.normal = m->dnormal;
  m->m.area_tab = m->darea_tab;
  m->m.kind = m->dkind;
  m->m.info = m->dinfo;
  m->m.vol = m->dvol;
  m->m.flags = nullptr;
  m->m.off = m->doff;
  m->m.inc = m->dinc;
  *out = m;
  return MMF_OK;
}

int mmf_mesh_sizes(const mmf_mesh *m, int64_t *ncells, int64_t *nfaces) {
  if (!m) return MMF_ERR_INVALID;
  if (ncells) *ncells = m->m.N;
  if (nfaces) *nfaces = m->m.F;
  return MMF_OK;
}

int mmf_mesh_set_cell_flags(mmf_mesh *m, const uint8_t *fluid, const uint8_t *solved) {
  if (!m) return MMF_ERR_INVALID;
  const int64_t N = m->m.N;
  std::vector<uint8_t> fl((size_t)N), bits((size_t)N);
  bool all = true;
  for (int64_t c = 0; c < N; ++c) {
    const uint8_t f = fluid ? (fluid[c] ? 1 : 0) : 1;
    const uint8_t s = solved ? (solved[c] ? 1 : 0) : f;
    if (s && !f) return mfail(m, MMF_ERR_INVALID, "a solved cell must be fluid (main.cpp:230-236)");
    fl[(size_t)c] = f;
    bits[(size_t)c] = (uint8_t)((f ? MMF_CELL_FLUID : 0) | (s ? MMF_CELL_SOLVED : 0));
    all = all && bits[(size_t)c] == (MMF_CELL_FLUID | MMF_CELL_SOLVED);
  }
  MCHK(m, hipStreamSynchronize(m->stream));
  int rc = upload_kinds(m, fl);
  if (rc != MMF_OK) return rc;
  drop_graphs(m);
  if (all) {
    if (m->dflags) (void)hipFree(m->dflags);
    m->dflags = nullptr;
    m->m.flags = nullptr;
    return MMF_OK;
  }
  if (!m->dflags) MCHK(m, hipMalloc(&m->dflags, (size_t)N));
  MCHK(m, hipMemcpy(m->dflags, bits.data(), (size_t)N, hipMemcpyHostToDevice));
  m->m.flags = m->dflags;
  return MMF_OK;
}

int mmf_mesh_set_state(mmf_mesh *m, const double *U) {
  if (!m || !U) return MMF_ERR_INVALID;
  const int64_t N = m->m.N, fs = m->m.fs;
  std::vector<double> soa((size_t)(5 * fs), 0.0);
  for (int64_t c = 0; c < N; ++c)
    for (int k = 0; k < 5; ++k) soa[(size_t)(k * fs + c)] = U[5 * c + k];
  MCHK(m, hipStreamSynchronize(m->stream));
  MCHK(m, hipMemcpy(m->dU, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
  // the stage buffer starts as a copy, so cells never solved hold finite values
  MCHK(m, hipMemcpy(m->dW, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
  return MMF_OK;
}

static int get_soa(mmf_mesh *m, const double *dsrc, double *U) {
  const int64_t N = m->m.N, fs = m->m.fs;
  std::vector<double> soa((size_t)(5 * fs));
  MCHK(m, hipStreamSynchronize(m->stream));
  MCHK(m, hipMemcpy(soa.data(), dsrc, sizeof(double) * soa.size(), hipMemcpyDeviceToHost));
  for (int64_t c = 0; c < N; ++c)
    for (int k = 0; k < 5; ++k) U[5 * c + k] = soa[(size_t)(k * fs + c)];
  return MMF_OK;
}

int mmf_mesh_get_state(mmf_mesh *m, double *U) {
  if (!m || !U) return MMF_ERR_INVALID;
  return get_soa(m, m->dU, U);
}

int mmf_mesh_compute_rhs(mmf_mesh *m, double *rhs, double *max_eig) {
  if (!m) return MMF_ERR_INVALID;
  hipError_t e = mmf::launch_zero_slot(m->dctrl, 3, m->stream);
  if (e != hipSuccess) return mfail(m, MMF_ERR_HIP, "zero slot");
  int rc = launch_stage(m, 1, mmf::MODE_RHS_ONLY);
  if (rc != MMF_OK) return rc;
  if (rhs) {
    rc = get_soa(m, m->dR, rhs);
    if (rc != MMF_OK) return rc;
  }
  rc = read_ctrl(m);
  if (rc != MMF_OK) return rc;
  if (max_eig) *max_eig = mmf::ctrl_stage_max(*m->hctrl, 3);
  return MMF_OK;
}

int mmf_mesh_set_time(mmf_mesh *m, double t, double tmax, double cfl) {
  if (!m || !(cfl > 0.0)) return mfail(m, MMF_ERR_INVALID, "cfl must be positive");
  int rc = read_ctrl(m);
  if (rc != MMF_OK) return rc;
  Ctrl c;
  std::memset(&c, 0, sizeof(c));
  c.t = t;
  c.tmax = tmax;
  c.cfl = cfl;
  c.hmin = m->hmin;  // main.cpp:351-356
  c.neg_p = m->hctrl->neg_p;
  c.neg_rho = m->hctrl->neg_rho;
  *m->hctrl = c;
  MCHK(m, hipMemcpyAsync(m->dctrl, m->hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return MMF_OK;
}

int mmf_mesh_advance(mmf_mesh *m, int64_t nsteps) {
  if (!m || nsteps < 0) return MMF_ERR_INVALID;
  if (m->timing || !m->use_graphs || nsteps == 0) return enqueue_steps(m, nsteps);
  while (nsteps > 0) {
    const int slot = (nsteps >= GRAPH_STEPS) ? 1 : 0;
    if (!m->exec[slot]) {
      int rc = capture(m, slot, slot ? GRAPH_STEPS : 1);
      if (rc != MMF_OK) {
        drop_graphs(m);
        return rc;
      }
    }
    MCHK(m, hipGraphLaunch(m->exec[slot], m->stream));
    nsteps -= slot ? GRAPH_STEPS : 1;
  }
  return MMF_OK;
}

int mmf_mesh_set_step_graphs(mmf_mesh *m, int32_t on) {
  if (!m) return MMF_ERR_INVALID;
  MCHK(m, hipStreamSynchronize(m->stream));
  m->use_graphs = on != 0;
  if (!m->use_graphs) drop_graphs(m);
  return MMF_OK;
}

int mmf_mesh_run(mmf_mesh *m, int64_t max_steps, int64_t *steps_done, double *t_final) {
  if (!m) return MMF_ERR_INVALID;
  int rc = read_ctrl(m);
  if (rc != MMF_OK) return rc;
  const int64_t step0 = m->hctrl->step;
  for (;;) {
    const Ctrl &c = *m->hctrl;
    const int64_t done = c.step - step0;
    if (c.nonfinite || !(c.t < c.tmax)) break;
    if (max_steps >= 0 && done >= max_steps) break;
    int64_t batch = 1;
    if (c.dt > 0.0) batch = (int64_t)std::min(std::max(std::ceil((c.tmax - c.t) / c.dt), 1.0), 256.0);
    if (max_steps >= 0) batch = std::min(batch, max_steps - done);
    rc = mmf_mesh_advance(m, batch);
    if (rc != MMF_OK) return rc;
    rc = read_ctrl(m);
    if (rc != MMF_OK) return rc;
  }
  if (steps_done) *steps_done = m->hctrl->step - step0;
  if (t_final) *t_final = m->hctrl->t;
  return nonfinite_status(m);
}

int mmf_mesh_get_time(mmf_mesh *m, double *t, double *dt, int64_t *step, double *stage_eig) {
  if (!m) return MMF_ERR_INVALID;
  int rc = read_ctrl(m);
  if (rc != MMF_OK) return rc;
  const Ctrl &c = *m->hctrl;
  if (t) *t = c.t;
  if (dt) *dt = c.dt;
  if (step) *step = c.step;
  if (stage_eig)
    for (int s = 0; s < 3; ++s) stage_eig[s] = c.last_eig[s];
  return nonfinite_status(m);
}

int mmf_mesh_get_counters(mmf_mesh *m, int64_t *neg_p, int64_t *neg_rho, int64_t *nonfinite_steps) {
  if (!m) return MMF_ERR_INVALID;
  int rc = read_ctrl(m);
  if (rc != MMF_OK) return rc;
  if (neg_p) *neg_p = (int64_t)m->hctrl->neg_p;
  if (neg_rho) *neg_rho = (int64_t)m->hctrl->neg_rho;
  if (nonfinite_steps) *nonfinite_steps = (int64_t)m->hctrl->nonfinite;
  return MMF_OK;
}

int mmf_mesh_enable_timing(mmf_mesh *m, int32_t on) {
  if (!m) return MMF_ERR_INVALID;
  int rc = collect_timing(m);
  if (rc != MMF_OK) return rc;
  m->timing = on != 0;
  m->face_ms = m->cell_ms = 0.0;
  m->stages = 0;
  return MMF_OK;
}

int mmf_mesh_read_timing(mmf_mesh *m, double *face_ms, double *cell_ms, int64_t *stages) {
  if (!m) return MMF_ERR_INVALID;
  int rc = collect_timing(m);
  if (rc != MMF_OK) return rc;
  if (face_ms) *face_ms = m->face_ms;
  if (cell_ms) *cell_ms = m->cell_ms;
  if (stages) *stages = m->stages;
  return MMF_OK;
}

}  // extern "C"

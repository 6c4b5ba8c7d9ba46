// Mechanism construction and finalization for griffon_amd: host-side tables and their device image.
//
// Setters follow griffon's (chemistry_setup.cpp:21-345); finalize() restates ReactionRateData::finalize
// (chemistry_setup.cpp:460-732: net species, stoichiometric sums, kForm and troeForm thresholds) and adds
// the gather tables this implementation's kernels use to assemble rates and Jacobians without atomics.
#include "gf_internal.h"
#include "../../include/griffon_amd.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <tuple>

namespace gfa {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

namespace {

struct Arena {
  std::vector<char> buf;
  template <typename T>
  size_t put(const std::vector<T> &v) {
    size_t off = (buf.size() + 255) & ~size_t(255);
    buf.resize(off + std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty()) std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

int finalize(gf_mech *m) {
  const int ns = (int)m->species.size();
  if (ns < 2) return fail(GF_ERR_STATE, "mechanism needs at least two species");
  if (ns > NS_LIMIT) return fail(GF_ERR_UNSUPPORTED, "more than 63 species is not supported");
  for (int i = 0; i < ns; ++i)
    if (m->species[i].cpkind == 0) return fail(GF_ERR_STATE, "species " + std::to_string(i) + " has no heat capacity data");
  const int nr = (int)m->rxns.size();
  std::vector<double> invmw(ns), mw(ns), dinvm(ns), Tmin(ns), Tmax(ns), spc(ns * 16);
  std::vector<int> cpkind(ns);
  for (int i = 0; i < ns; ++i) {
    mw[i] = m->species[i].mw;
    invmw[i] = 1. / mw[i];
    Tmin[i] = m->species[i].Tmin;
    Tmax[i] = m->species[i].Tmax;
    cpkind[i] = m->species[i].cpkind;
    for (int k = 0; k < 16; ++k) spc[i * 16 + k] = m->species[i].c[k];
  }
  for (int i = 0; i < ns; ++i) dinvm[i] = invmw[i] - invmw[ns - 1];

  std::vector<DevRxn> drx(nr);
  std::vector<int> tb_idx, tb_slot, sp_idx, slot_sp;
  std::vector<double> tb_eff, sp_ord, sp_invmw;
  std::vector<std::vector<int>> slot_lists(nr);

  for (int r = 0; r < nr; ++r) {
    const HostRxn &h = m->rxns[r];
    DevRxn &d = drx[r];
    std::memset(&d, 0, sizeof(d));
    d.type = h.type;
    d.hasOrders = h.sp_idx.empty() ? 0 : 1;
    d.reversible = (h.reversible && !d.hasOrders) ? 1 : 0;  // special orders imply kr = 0 (chemistry_kernels.cpp:325-336)
    for (int k = 0; k < 3; ++k) { d.kf[k] = h.kf[k]; d.kp[k] = h.kp[k]; }
    for (int k = 0; k < 4; ++k) d.troe[k] = h.troe[k];
    d.baseEff = h.default_eff;
    d.n_rc = (int)h.rc_idx.size();
    d.n_pd = (int)h.pd_idx.size();
    if (d.n_rc > NRC_MAX || d.n_pd > NPD_MAX) return fail(GF_ERR_UNSUPPORTED, "more than 4 reactants or products");
    // net species (chemistry_setup.cpp:463-541), ordered by species index
    std::map<int, int> net;
    d.sumStoich = 0;
    for (int j = 0; j < d.n_rc; ++j) { net[h.rc_idx[j]] += h.rc_st[j]; d.sumStoich += h.rc_st[j]; }
    for (int j = 0; j < d.n_pd; ++j) { net[h.pd_idx[j]] -= h.pd_st[j]; d.sumStoich -= h.pd_st[j]; }
    int nn = 0;
    for (auto &kv : net) if (kv.second != 0) ++nn;
    if (nn < 2 || nn > NNET_MAX)
      return fail(GF_ERR_ARG, "reaction " + std::to_string(r) + ": bad number of net reacting species (n < 2 or n > 8)");
    d.n_net = 0;
    for (auto &kv : net)
      if (kv.second != 0) { d.net_idx[d.n_net] = (uint8_t)kv.first; d.net_st[d.n_net] = (int8_t)kv.second; ++d.n_net; }
    d.sumRc = 0;
    for (int j = 0; j < d.n_rc; ++j) { d.rc_idx[j] = (uint8_t)h.rc_idx[j]; d.rc_st[j] = (uint8_t)h.rc_st[j]; d.sumRc += h.rc_st[j]; }
    d.sumPd = 0;
    for (int j = 0; j < d.n_pd; ++j) { d.pd_idx[j] = (uint8_t)h.pd_idx[j]; d.pd_st[j] = (uint8_t)h.pd_st[j]; d.sumPd += h.pd_st[j]; }
    // kForm (chemistry_setup.cpp:674-688)
    if (std::abs(h.kf[2]) < 1.e-6) {
      if (std::abs(h.kf[1]) < 1.e-6) d.kForm = KF_CONSTANT;
      else if (std::abs(h.kf[1] - 1) < 1.e-6) d.kForm = KF_LINEAR;
      else if (std::abs(h.kf[1] - 2) < 1.e-6) d.kForm = KF_QUADRATIC;
      else if (std::abs(h.kf[1] + 1) < 1.e-6) d.kForm = KF_RECIPROCAL;
      else d.kForm = KF_ARRHENIUS;
    } else {
      d.kForm = KF_ARRHENIUS;
    }
    // troeForm (chemistry_setup.cpp:690-730)
    d.troeForm = TF_NONE;
    if (h.type == RK_TROE) {
      const bool t1 = std::abs(h.troe[1]) > 1.e-8, t2 = std::abs(h.troe[2]) > 1.e-8, t3 = std::abs(h.troe[3]) > 1.e-8;
      if (t1) d.troeForm = t2 ? (t3 ? TF_T123 : TF_T12) : (t3 ? TF_T13 : TF_T1);
      else d.troeForm = t2 ? (t3 ? TF_T23 : TF_T2) : (t3 ? TF_T3 : TF_NONE);
      if (d.troeForm == TF_NONE) return fail(GF_ERR_ARG, "Troe reaction without Troe terms");
    }
    // slots: distinct participating species other than the last one
    std::vector<int> &sl = slot_lists[r];
    auto slot_of = [&](int sp) -> int {
      if (sp == ns - 1) return -1;
      for (size_t q = 0; q < sl.size(); ++q) if (sl[q] == sp) return (int)q;
      sl.push_back(sp);
      return (int)sl.size() - 1;
    };
    for (int j = 0; j < d.n_rc; ++j) d.rc_slot[j] = (int8_t)slot_of(h.rc_idx[j]);
    for (int j = 0; j < d.n_pd; ++j) d.pd_slot[j] = (int8_t)slot_of(h.pd_idx[j]);
    d.tb_off = (int)tb_idx.size();
    d.n_tb = 0;
    d.tb_last = -1;
    if (h.type != RK_SIMPLE) {
      for (size_t j = 0; j < h.tb_idx.size(); ++j) {
        const int sp = h.tb_idx[j];
        if (sp == ns - 1 && d.tb_last < 0) d.tb_last = (int8_t)d.n_tb;
        tb_idx.push_back(sp);
        tb_eff.push_back(invmw[sp] * (h.tb_eff_raw[j] - h.default_eff));
        tb_slot.push_back(slot_of(sp));
        ++d.n_tb;
      }
    }
    d.sp_off = (int)sp_idx.size();
    d.n_sp = 0;
    for (size_t j = 0; j < h.sp_idx.size(); ++j) {
      if (!(std::abs(h.sp_ord[j]) > 1.e-12)) continue;  // zero orders are skipped (chemistry_setup.cpp:431)
      if (d.n_sp == NSP_MAX) return fail(GF_ERR_UNSUPPORTED, "more than 4 species with special orders");
      sp_idx.push_back(h.sp_idx[j]);
      sp_ord.push_back(h.sp_ord[j]);
      sp_invmw.push_back(invmw[h.sp_idx[j]]);
      d.sp_slot[d.n_sp] = (int8_t)slot_of(h.sp_idx[j]);
      ++d.n_sp;
    }
    if ((int)sl.size() > 127) return fail(GF_ERR_UNSUPPORTED, "too many participants in one reaction");
    d.slot_off = (int)slot_sp.size();
    d.n_slot = (int)sl.size();
    for (int sp : sl) slot_sp.push_back(sp);
  }
  const int nslot = (int)slot_sp.size();

  const int nm = ns - 1;
  // species <- reactions (net stoichiometry) with the slot of that species in the reaction, if any
  std::vector<uint32_t> sgx_off(ns + 1, 0);
  std::vector<uint64_t> sgx;
  {
    std::vector<std::vector<uint64_t>> L(ns);
    for (int r = 0; r < nr; ++r) {
      const DevRxn &d = drx[r];
      for (int q = 0; q < d.n_net; ++q) {
        const int k = d.net_idx[q];
        uint64_t dslot = 0xFFFF;
        for (int p = 0; p < d.n_slot; ++p)
          if (slot_sp[d.slot_off + p] == k) dslot = (uint64_t)(d.slot_off + p);
        L[k].push_back((uint64_t)r | (dslot << 16) | ((uint64_t)(uint8_t)d.net_st[q] << 32));
      }
    }
    for (int k = 0; k < ns; ++k) {
      sgx_off[k + 1] = sgx_off[k] + (uint32_t)L[k].size();
      sgx.insert(sgx.end(), L[k].begin(), L[k].end());
    }
  }
  // eg: all terms of raw sensitivity entry (i, j); exx: the off-diagonal ones in compact form (the
  // diagonal terms are summed in the species gather above)
  std::vector<int> eg_off(nm * ns + 1, 0), eg_slot;
  std::vector<double> eg_coef;
  std::vector<uint16_t> exx_off(nm * ns + 1, 0);
  std::vector<uint32_t> exx;
  {
    std::vector<std::vector<std::pair<int, int>>> E(nm * ns);  // (slot, nu of the row species)
    for (int r = 0; r < nr; ++r) {
      const DevRxn &d = drx[r];
      for (int p = 0; p < d.n_slot; ++p) {
        const int j = slot_sp[d.slot_off + p];
        for (int q = 0; q < d.n_net; ++q) E[j * ns + d.net_idx[q]].push_back({d.slot_off + p, d.net_st[q]});
      }
    }
    for (int e = 0; e < nm * ns; ++e) {
      const int i = e % ns;
      eg_off[e + 1] = eg_off[e] + (int)E[e].size();
      for (auto &t : E[e]) { eg_slot.push_back(t.first); eg_coef.push_back(-t.second * mw[i]); }
      const bool diag = (e / ns) == i;
      if (!diag)
        for (auto &t : E[e]) exx.push_back((uint32_t)t.first | ((uint32_t)(uint8_t)(int8_t)t.second << 16));
      if (exx.size() > 0xFFFF) return fail(GF_ERR_UNSUPPORTED, "too many Jacobian sensitivity terms");
      exx_off[e + 1] = (uint16_t)exx.size();
    }
  }
  if (nslot > 0xFFFE || nr > 0xFFFF || sgx.size() > 0xFFFFFFu)
    return fail(GF_ERR_UNSUPPORTED, "mechanism too large for the compact gather tables");
  // Device order of the reactions: the rate kernels give a reaction (times a few states) to each lane, so
  // reactions of one kind (special orders, type, Troe form, reversibility, rate form, reactant/product/third-body
  // counts) are stored together and a wave takes one branch. Every table above was built in mechanism order
  // and keeps that order (sums over reactions run as before); only the reaction records move, each keeping its
  // mechanism index in `orig`, and the species gather is re-pointed at the new positions.
  {
    std::vector<int> perm(nr);
    for (int r = 0; r < nr; ++r) perm[r] = r;
    auto key = [&](int r) {
      const DevRxn &d = drx[r];
      return std::make_tuple(d.hasOrders, d.type, d.troeForm, d.reversible, d.kForm, d.n_rc, d.n_pd, d.n_tb);
    };
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key(a) < key(b); });
    std::vector<int> pos(nr);
    std::vector<DevRxn> sorted(nr);
    for (int d = 0; d < nr; ++d) {
      pos[perm[d]] = d;
      sorted[d] = drx[perm[d]];
      sorted[d].orig = perm[d];
    }
    drx.swap(sorted);
    for (auto &e : sgx) e = (e & ~uint64_t(0xFFFF)) | (uint64_t)pos[e & 0xFFFFu];
  }
  // byte image of the LDS tables
  std::vector<char> tbl;
  auto tput = [&](const void *src, size_t bytes) {
    const size_t at = (tbl.size() + 7) & ~size_t(7);
    tbl.resize(at + bytes);
    if (bytes) memcpy(tbl.data() + at, src, bytes);
    return (int)at;
  };
  tput(sgx_off.data(), sgx_off.size() * sizeof(uint32_t));
  const int sgx_at = tput(sgx.data(), sgx.size() * sizeof(uint64_t));
  const int tbe_at = tput(tb_eff.data(), tb_eff.size() * sizeof(double));
  const int tbi_at = tput(tb_idx.data(), tb_idx.size() * sizeof(int));
  const int tbs_at = tput(tb_slot.data(), tb_slot.size() * sizeof(int));
  const int imw_at = tput(invmw.data(), invmw.size() * sizeof(double));
  const int exo_at = tput(exx_off.data(), exx_off.size() * sizeof(uint16_t));
  const int exx_at = tput(exx.data(), exx.size() * sizeof(uint32_t));
  tbl.resize((tbl.size() + 15) & ~size_t(15));

  Arena a;
  MechView &v = m->view_host;
  v.ns = ns; v.nr = nr; v.nslot = nslot;
  v.last_const = -1;
  for (int i = 0; i < ns; ++i) if (cpkind[i] == CP_CONST) v.last_const = i;
  v.Ru = m->Ru; v.invRu = 1. / m->Ru; v.pref = m->pref;
  auto P = [](size_t off) { return reinterpret_cast<const char *>(off); };
  v.spc_c = (const double *)P(a.put(spc));
  v.invmw = (const double *)P(a.put(invmw));
  v.mw = (const double *)P(a.put(mw));
  v.dinvm = (const double *)P(a.put(dinvm));
  v.Tmin = (const double *)P(a.put(Tmin));
  v.Tmax = (const double *)P(a.put(Tmax));
  v.cpkind = (const int *)P(a.put(cpkind));
  v.rx = (const DevRxn *)P(a.put(drx));
  v.tb_idx = (const int *)P(a.put(tb_idx));
  v.tb_eff = (const double *)P(a.put(tb_eff));
  v.tb_slot = (const int *)P(a.put(tb_slot));
  v.sp_idx = (const int *)P(a.put(sp_idx));
  v.sp_ord = (const double *)P(a.put(sp_ord));
  v.sp_invmw = (const double *)P(a.put(sp_invmw));
  v.eg_off = (const int *)P(a.put(eg_off));
  v.eg_slot = (const int *)P(a.put(eg_slot));
  v.eg_coef = (const double *)P(a.put(eg_coef));
  v.tbl = (const char *)P(a.put(tbl));
  v.tbl_bytes = (int)tbl.size();
  v.sgx_at = sgx_at; v.exo_at = exo_at; v.exx_at = exx_at;
  v.tbe_at = tbe_at; v.tbi_at = tbi_at; v.tbs_at = tbs_at; v.imw_at = imw_at;
  m->blob.swap(a.buf);
  m->drx = drx;
  m->lds_per_state = sizeof(double) * (12 * NS_MAX + 16 + 5 * (size_t)nr + (size_t)nslot);
  m->lds_per_state_rhs = sizeof(double) * (4 * NS_MAX + 16 + (size_t)nr);
  m->finalized = true;
  return GF_OK;
}

template <typename T>
const T *rebase(const T *off, const char *base) {
  return reinterpret_cast<const T *>(base + reinterpret_cast<size_t>(off));
}

}  // namespace

int ensure_device(gf_mech *m) {
  if (!m->finalized) {
    const int rc = finalize(m);
    if (rc != GF_OK) return rc;
  }
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return fail(GF_ERR_HIP, "no HIP device available");
  if (m->dblob && m->device == dev) return GF_OK;
  if (m->dblob) {
    (void)hipFree(m->dblob);
    m->dblob = nullptr;
  }
  if (hipMalloc(&m->dblob, m->blob.size()) != hipSuccess) return fail(GF_ERR_HIP, "hipMalloc of mechanism tables failed");
  if (hipMemcpy(m->dblob, m->blob.data(), m->blob.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail(GF_ERR_HIP, "upload of mechanism tables failed");
  m->device = dev;
  int ncu = 0, lmax = 0, lopt = 0;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  (void)hipDeviceGetAttribute(&lopt, hipDeviceAttributeSharedMemPerBlockOptin, dev);
  m->n_cu = ncu > 0 ? ncu : 1;
  m->lds_max = std::max(lmax, lopt) > 0 ? (size_t)std::max(lmax, lopt) : 65536;
  const char *b = static_cast<const char *>(m->dblob);
  MechView v = m->view_host;
  v.spc_c = rebase(v.spc_c, b); v.invmw = rebase(v.invmw, b); v.mw = rebase(v.mw, b); v.dinvm = rebase(v.dinvm, b);
  v.Tmin = rebase(v.Tmin, b); v.Tmax = rebase(v.Tmax, b); v.cpkind = rebase(v.cpkind, b); v.rx = rebase(v.rx, b);
  v.tb_idx = rebase(v.tb_idx, b); v.tb_eff = rebase(v.tb_eff, b); v.tb_slot = rebase(v.tb_slot, b);
  v.sp_idx = rebase(v.sp_idx, b); v.sp_ord = rebase(v.sp_ord, b); v.sp_invmw = rebase(v.sp_invmw, b);
  v.eg_off = rebase(v.eg_off, b); v.eg_slot = rebase(v.eg_slot, b); v.eg_coef = rebase(v.eg_coef, b);
  v.tbl = rebase(v.tbl, b);
  m->view = v;
  if (!m->stream) {
    if (hipStreamCreate(&m->stream) != hipSuccess) return fail(GF_ERR_HIP, "hipStreamCreate failed");
  }
  return GF_OK;
}

int ensure_stage(gf_mech *m, size_t bytes) {
  if (m->stage_bytes >= bytes) return GF_OK;
  if (m->stage) (void)hipFree(m->stage);
  m->stage = nullptr;
  m->stage_bytes = 0;
  if (hipMalloc(&m->stage, bytes) != hipSuccess) return fail(GF_ERR_HIP, "hipMalloc of staging buffer failed");
  m->stage_bytes = bytes;
  return GF_OK;
}

int ensure_work(gf_mech *m, size_t bytes) {
  if (m->work_bytes >= bytes) return GF_OK;
  if (m->work) (void)hipFree(m->work);
  m->work = nullptr;
  m->work_bytes = 0;
  if (hipMalloc(&m->work, bytes) != hipSuccess) return fail(GF_ERR_HIP, "hipMalloc of the workspace failed");
  m->work_bytes = bytes;
  return GF_OK;
}

}  // namespace gfa

using namespace gfa;

extern "C" {

const char *gf_last_error(void) { return g_last_error.c_str(); }

int gf_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int gf_mech_create(gf_mech **out) {
  if (!out) return fail(GF_ERR_ARG, "null output pointer");
  *out = new gf_mech();
  return GF_OK;
}

int gf_mech_destroy(gf_mech *m) {
  if (!m) return GF_OK;
  if (m->dblob) (void)hipFree(m->dblob);
  if (m->stage) (void)hipFree(m->stage);
  if (m->work) (void)hipFree(m->work);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return GF_OK;
}

#define GF_CHECK_BUILDING(m)                                                     \
  do {                                                                           \
    if (!(m)) return fail(GF_ERR_ARG, "null mechanism");                         \
    if ((m)->finalized) return fail(GF_ERR_STATE, "mechanism already finalized"); \
  } while (0)

int gf_mechanism_set_gas_constant(gf_mech *m, double Ru) { GF_CHECK_BUILDING(m); m->Ru = Ru; return GF_OK; }
int gf_mechanism_set_ref_pressure(gf_mech *m, double p) { GF_CHECK_BUILDING(m); m->pref = p; return GF_OK; }
int gf_mechanism_set_ref_temperature(gf_mech *m, double T) { GF_CHECK_BUILDING(m); m->Tref = T; return GF_OK; }

int gf_mechanism_add_species(gf_mech *m, double mw) {
  GF_CHECK_BUILDING(m);
  if (!(mw > 0.)) return fail(GF_ERR_ARG, "molecular weight must be positive");
  Species s;
  s.mw = mw;
  m->species.push_back(s);
  return GF_OK;
}

int gf_mechanism_resize_heat_capacity_data(gf_mech *m) { GF_CHECK_BUILDING(m); return GF_OK; }

int gf_mechanism_add_const_cp(gf_mech *m, int i, double Tmin, double Tmax, double T0, double h0, double s0, double cp) {
  GF_CHECK_BUILDING(m);
  if (i < 0 || i >= (int)m->species.size()) return fail(GF_ERR_ARG, "species index out of range");
  Species &s = m->species[i];
  s.cpkind = CP_CONST;
  s.Tmin = Tmin; s.Tmax = Tmax;
  std::memset(s.c, 0, sizeof(s.c));
  s.c[0] = T0; s.c[1] = h0; s.c[2] = s0; s.c[3] = cp;
  return GF_OK;
}

// chemistry_setup.cpp:102-130: coefficients scaled by R and pre-divided by 2, 6, 12, 20
int gf_mechanism_add_nasa7_cp(gf_mech *m, int i, double Tmin, double Tmid, double Tmax, const double *low,
                              const double *high) {
  GF_CHECK_BUILDING(m);
  if (i < 0 || i >= (int)m->species.size()) return fail(GF_ERR_ARG, "species index out of range");
  if (!low || !high) return fail(GF_ERR_ARG, "null coefficient array");
  Species &s = m->species[i];
  s.cpkind = CP_NASA7;
  s.Tmin = Tmin; s.Tmax = Tmax;
  std::memset(s.c, 0, sizeof(s.c));
  s.c[0] = Tmid;
  for (int k = 0; k < 7; ++k) s.c[1 + k] = high[k] * m->Ru;
  for (int k = 0; k < 7; ++k) s.c[8 + k] = low[k] * m->Ru;
  s.c[2] /= 2.; s.c[3] /= 6.; s.c[4] /= 12.; s.c[5] /= 20.;
  s.c[9] /= 2.; s.c[10] /= 6.; s.c[11] /= 12.; s.c[12] /= 20.;
  return GF_OK;
}

int gf_mechanism_add_reaction(gf_mech *m, int type, int reversible, int n_r, const int *r_idx, const int *r_st, int n_p,
                              const int *p_idx, const int *p_st, const double *kfwd, int n_tb, const int *tb_idx,
                              const double *tb_eff, double default_eff, const double *kpres, const double *troe,
                              int n_ord, const int *ord_idx, const double *ord_val) {
  GF_CHECK_BUILDING(m);
  const int ns = (int)m->species.size();
  if (type < RK_SIMPLE || type > RK_TROE) return fail(GF_ERR_ARG, "unknown reaction type");
  if (n_r < 0 || n_p < 0 || n_tb < 0 || n_ord < 0 || !kfwd) return fail(GF_ERR_ARG, "bad reaction arguments");
  HostRxn h;
  h.type = type;
  h.reversible = reversible != 0;
  auto chk = [&](int idx) { return idx >= 0 && idx < ns; };
  for (int j = 0; j < n_r; ++j) {
    if (!chk(r_idx[j]) || r_st[j] < 1 || r_st[j] > 3) return fail(GF_ERR_ARG, "bad reactant (index or stoichiometry outside 1..3)");
    h.rc_idx.push_back(r_idx[j]); h.rc_st.push_back(r_st[j]);
  }
  for (int j = 0; j < n_p; ++j) {
    if (!chk(p_idx[j]) || p_st[j] < 1 || p_st[j] > 3) return fail(GF_ERR_ARG, "bad product (index or stoichiometry outside 1..3)");
    h.pd_idx.push_back(p_idx[j]); h.pd_st.push_back(p_st[j]);
  }
  for (int k = 0; k < 3; ++k) h.kf[k] = kfwd[k];
  if (type != RK_SIMPLE) {
    h.default_eff = default_eff;
    for (int j = 0; j < n_tb; ++j) {
      if (!chk(tb_idx[j])) return fail(GF_ERR_ARG, "bad third-body index");
      h.tb_idx.push_back(tb_idx[j]); h.tb_eff_raw.push_back(tb_eff[j]);
    }
  }
  if (type == RK_LINDEMANN || type == RK_TROE) {
    if (!kpres) return fail(GF_ERR_ARG, "falloff reaction without low-pressure rate");
    for (int k = 0; k < 3; ++k) h.kp[k] = kpres[k];
  }
  if (type == RK_TROE) {
    if (!troe) return fail(GF_ERR_ARG, "Troe reaction without parameters");
    for (int k = 0; k < 4; ++k) h.troe[k] = troe[k];
  }
  for (int j = 0; j < n_ord; ++j) {
    if (!chk(ord_idx[j])) return fail(GF_ERR_ARG, "bad special-order index");
    h.sp_idx.push_back(ord_idx[j]); h.sp_ord.push_back(ord_val[j]);
  }
  m->rxns.push_back(h);
  return GF_OK;
}

int gf_mech_n_species(const gf_mech *m) { return m ? (int)m->species.size() : 0; }
int gf_mech_n_reactions(const gf_mech *m) { return m ? (int)m->rxns.size() : 0; }

}  // extern "C"

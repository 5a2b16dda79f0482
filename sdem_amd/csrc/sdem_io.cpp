// sdem_io.cpp -- the reference's restart and visualisation files for one
// box of a context (include/sdem_io.h).  Host code over the public entry
// points (sdem_get_state, sdem_get_contacts, sdem_set_state, ...) and the
// internal box description (sdem_host.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/sdem.h"
#include "../../include/sdem_io.h"
#include "sdem_host.h"

namespace {

constexpr int MAX_LINKS = 20;  // CntctVec(20) / "so big ncv" (Prtcl_CL_and_CF.f90:641-680)

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) { f = std::fopen(path, mode); }
  ~File() {
    if (f) std::fclose(f);
  }
  bool put(const void* p, size_t bytes) { return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes; }
  bool get(void* p, size_t bytes) { return bytes == 0 || std::fread(p, 1, bytes, f) == bytes; }
};

int fail(sdem_ctx* c, int code, const std::string& m) { return sdem_internal_fail(c, code, m.c_str()); }

// the state of the selected box in ascending id order, in the reference's
// per-particle layouts
struct State {
  int n = 0, ts = 1, rs = 1;
  std::vector<int32_t> id, type, mark;
  std::vector<double> posr, lv, la, th, rv, ra;
};

int read_state(sdem_ctx* c, const sdem_box_info& I, State& S) {
  int32_t n = 0;
  if (int rc = sdem_get_num_particles(c, &n)) return rc;
  S.n = n;
  S.ts = I.tsize;
  S.rs = I.rsize;
  S.id.resize(n); S.type.resize(n); S.mark.resize(n);
  S.posr.resize(4 * (size_t)n);
  S.lv.resize(3 * (size_t)S.ts * n); S.la.resize(3 * (size_t)S.ts * n);
  S.th.resize(3 * (size_t)n);
  S.rv.resize(3 * (size_t)S.rs * n); S.ra.resize(3 * (size_t)S.rs * n);
  int32_t got = 0;
  return sdem_get_state(c, n, &got, S.id.data(), S.type.data(), S.mark.data(), S.posr.data(), S.lv.data(),
                        S.la.data(), S.th.data(), S.rv.data(), S.ra.data());
}

// Fortran integer edit descriptors: Iw right-justified in w columns, Iw.w
// zero-padded to w digits; a value that does not fit prints w asterisks
std::string fI(long long v, int w, bool zero_pad = false) {
  char buf[64];
  if (zero_pad) std::snprintf(buf, sizeof buf, "%0*lld", w, v);
  else std::snprintf(buf, sizeof buf, "%*lld", w, v);
  std::string s(buf);
  if ((int)s.size() > w) s.assign(w, '*');
  return s;
}

std::string sp(int n) { return std::string(n, ' '); }

struct VisuField {
  uint32_t bit;
  const char *name, *attr, *dtype;
  int dims, prec;
};
// PIO_Write_XDMF's attribute order (Prtcl_IOAndVisu.f90:1086-1129)
const VisuField VISU_FIELDS[] = {
    {SDEM_VISU_ID, "ID", "Scalar", "Int", 1, 4},
    {SDEM_VISU_DIAMETER, "Diameter", "Scalar", "Float", 1, 8},
    {SDEM_VISU_TYPE, "Type", "Scalar", "Int", 1, 4},
    {SDEM_VISU_USRMARK, "UsrMark", "Scalar", "Int", 1, 4},
    {SDEM_VISU_LINVEL, "LinVel", "Vector", "Float", 3, 8},
    {SDEM_VISU_LINACC, "LinAcc", "Vector", "Float", 3, 8},
    {SDEM_VISU_THETA, "Theta", "Vector", "Float", 3, 8},
    {SDEM_VISU_ROTVEL, "RotVel", "Vector", "Float", 3, 8},
    {SDEM_VISU_ROTACC, "RotAcc", "Vector", "Float", 3, 8},
    {SDEM_VISU_CNTCTFORCE, "CntctForce", "Vector", "Float", 3, 8},
    {SDEM_VISU_TORQUE, "Torque", "Vector", "Float", 3, 8},
};
constexpr uint32_t VISU_ALL = (1u << 11) - 1u;

}  // namespace

// ---------------------------------------------------------------------------
// restart file
// ---------------------------------------------------------------------------

// PIO_Write_Restart (Prtcl_IOAndVisu.f90:849-1031).  The links of a particle
// are its own records (CL_Gather_Cntctlink_Restart walks Bucket first,
// Prtcl_CL_and_CF.f90:656-685) and then the PP records owned by a lower-id
// partner (the Head_Cp counterpart list); CLid numbers the records once
// each (CL_Prepare_Restart :344-359), and the TanDel block holds them in
// CLid order (CL_GetNextTanDel_Un).  Here records are numbered in
// (owner id, partner id) order and links follow that order.
extern "C" int sdem_write_restart(sdem_ctx* c, const char* path, int32_t mark_links) {
  if (!c || !path) return SDEM_E_ARG;
  sdem_box_info I;
  if (int rc = sdem_internal_box_info(c, &I)) return rc;
  State S;
  if (int rc = read_state(c, I, S)) return rc;
  if (S.n < 1) return SDEM_OK;  // np_InDomain < 1: no file (:869)
  int64_t m = 0;
  if (int rc = sdem_get_num_contacts(c, &m)) return rc;
  std::vector<int32_t> kind(m), owner(m), partner(m);
  std::vector<double> tan(3 * (size_t)m), vri(m);
  int64_t got = 0;
  if (m > 0)
    if (int rc = sdem_get_contacts(c, m, &got, kind.data(), owner.data(), partner.data(), tan.data(), vri.data()))
      return rc;
  const int n = S.n;
  std::unordered_map<int32_t, int> row;
  row.reserve(2 * (size_t)n);
  for (int i = 0; i < n; i++) row[S.id[i]] = i;
  // links per particle: own records, then counterpart (PP, partner side)
  std::vector<std::vector<std::pair<int32_t, int32_t>>> own(n), cp(n);
  for (int64_t r = 0; r < m; r++) {
    auto it = row.find(owner[r]);
    if (it == row.end()) return fail(c, SDEM_E_STATE, "contact record of an unknown particle");
    const int32_t clid = (int32_t)(r + 1);
    own[it->second].push_back({partner[r], clid});
    if (kind[r] == 1) {
      auto jt = row.find(partner[r]);
      if (jt == row.end()) return fail(c, SDEM_E_STATE, "PP record with an unknown partner");
      cp[jt->second].push_back({owner[r], clid});
    }
  }
  std::vector<int32_t> ncv(n);
  int64_t nlinks = 0;
  for (int i = 0; i < n; i++) {
    ncv[i] = (int32_t)(own[i].size() + cp[i].size());
    if (ncv[i] > MAX_LINKS)
      return fail(c, SDEM_E_CAPACITY,
                  "particle " + std::to_string(S.id[i]) + " has more than 20 contact links (CL_Count_Cntctlink)");
    nlinks += ncv[i];
  }
  if (nlinks > INT32_MAX) return fail(c, SDEM_E_CAPACITY, "too many contact links for the restart format");

  File F(path, "wb");
  if (!F.f) return fail(c, SDEM_E_ARG, std::string("cannot open ") + path);
  bool ok = true;
  const int32_t hdr[2] = {n, (int32_t)nlinks};
  ok = ok && F.put(hdr, sizeof hdr);
  {
    std::vector<double> p(3 * (size_t)n);
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) p[3 * (size_t)i + d] = S.posr[4 * (size_t)i + d];
    ok = ok && F.put(p.data(), p.size() * 8);
  }
  {
    std::vector<int32_t> im(3 * (size_t)n);
    for (int i = 0; i < n; i++) {
      im[3 * (size_t)i] = S.id[i];
      im[3 * (size_t)i + 1] = S.type[i];
      im[3 * (size_t)i + 2] = S.mark[i];
    }
    ok = ok && F.put(im.data(), im.size() * 4);
  }
  {
    const int ts = S.ts, rs = S.rs, nreal3 = 2 * ts + 2 * rs + 1;
    std::vector<double> blk(3 * (size_t)nreal3);
    for (int i = 0; i < n && ok; i++) {
      double* q = blk.data();
      auto cp3 = [&](const double* src, int cnt) {
        std::memcpy(q, src, 3 * sizeof(double) * cnt);
        q += 3 * cnt;
      };
      cp3(&S.lv[3 * (size_t)ts * i], ts);
      cp3(&S.la[3 * (size_t)ts * i], ts);
      cp3(&S.th[3 * (size_t)i], 1);
      cp3(&S.rv[3 * (size_t)rs * i], rs);
      cp3(&S.ra[3 * (size_t)rs * i], rs);
      ok = F.put(blk.data(), blk.size() * 8);
    }
  }
  ok = ok && F.put(ncv.data(), ncv.size() * 4);
  {
    std::vector<int32_t> lk;
    lk.reserve(2 * (size_t)nlinks);
    for (int i = 0; i < n; i++) {
      for (auto& l : own[i]) { lk.push_back(l.first); lk.push_back(l.second); }
      for (auto& l : cp[i]) { lk.push_back(l.first); lk.push_back(l.second); }
    }
    ok = ok && F.put(lk.data(), lk.size() * 4);
  }
  {
    std::vector<double> td(4 * (size_t)m);
    for (int64_t r = 0; r < m; r++) {
      td[4 * r] = tan[3 * r];
      td[4 * r + 1] = tan[3 * r + 1];
      td[4 * r + 2] = tan[3 * r + 2];
      td[4 * r + 3] = vri[r];
    }
    ok = ok && F.put(td.data(), td.size() * 8);
  }
  if (!ok) return fail(c, SDEM_E_ARG, std::string("write failed: ") + path);
  if (mark_links) return sdem_internal_set_marks(c, n, S.id.data(), ncv.data());
  return SDEM_OK;
}

// PIO_Read_Restart (Prtcl_IOAndVisu.f90:731-844) and PIO_RestartCL
// (:296-377) with CL_Add_RestartCntctlink (Prtcl_CL_and_CF.f90:577-622).
// A link (gid2, CLid) of particle gid1 becomes:
//   * nothing when gid2 is a moving particle with gid2 < gid1 within
//     1.5 (r1 + r2) (its record sits with gid2);
//   * a PP record of gid1 when gid2 > gid1 is a moving particle and the
//     pair is not across a periodic boundary, a PPG record of gid1 when it
//     is (each side keeps its own);
//   * a PPFix record for a fixed particle's id, a PW record for a wall id.
// A link to a lower-id particle that is neither close nor across a
// boundary is dropped: the reference keeps it in gid1's bucket, where the
// search never looks for that pair, and releases it after the first step.
extern "C" int sdem_read_restart(sdem_ctx* c, const char* path) {
  if (!c || !path) return SDEM_E_ARG;
  sdem_box_info I;
  if (int rc = sdem_internal_box_info(c, &I)) return rc;
  File F(path, "rb");
  if (!F.f) return fail(c, SDEM_E_ARG, std::string("cannot open file: ") + path);
  std::fseek(F.f, 0, SEEK_END);
  const long long fsize = std::ftell(F.f);
  std::fseek(F.f, 0, SEEK_SET);
  int32_t hdr[2];
  if (!F.get(hdr, sizeof hdr)) return fail(c, SDEM_E_ARG, "restart file too short");
  const int np = hdr[0];
  const int64_t nlinks = hdr[1];
  if (np < 0 || nlinks < 0) return fail(c, SDEM_E_ARG, "bad restart header");
  if (np > I.num_prtcl) return fail(c, SDEM_E_CAPACITY, "np_InDomain > numPrtcl");
  const int ts = I.tsize, rs = I.rsize, nreal3 = 2 * ts + 2 * rs + 1;
  const long long fixed_part = 8LL + (long long)np * (24 + 12 + 24LL * nreal3 + 4) + 8LL * nlinks;
  if (fsize < fixed_part || (fsize - fixed_part) % 32 != 0)
    return fail(c, SDEM_E_ARG, "restart file size does not match np, nLinks and the integrators");
  const int64_t ntan = (fsize - fixed_part) / 32;

  std::vector<double> pos(3 * (size_t)np);
  std::vector<int32_t> im(3 * (size_t)np), ncv(np), lk(2 * (size_t)nlinks);
  std::vector<double> blk(3 * (size_t)nreal3 * np), td(4 * (size_t)ntan);
  bool ok = F.get(pos.data(), pos.size() * 8) && F.get(im.data(), im.size() * 4) &&
            F.get(blk.data(), blk.size() * 8) && F.get(ncv.data(), ncv.size() * 4) &&
            F.get(lk.data(), lk.size() * 4) && F.get(td.data(), td.size() * 8);
  if (!ok) return fail(c, SDEM_E_ARG, "restart file too short");

  State S;
  S.n = np; S.ts = ts; S.rs = rs;
  S.id.resize(np); S.type.resize(np); S.mark.resize(np);
  S.posr.resize(4 * (size_t)np);
  S.lv.resize(3 * (size_t)ts * np); S.la.resize(3 * (size_t)ts * np); S.th.resize(3 * (size_t)np);
  S.rv.resize(3 * (size_t)rs * np); S.ra.resize(3 * (size_t)rs * np);
  for (int i = 0; i < np; i++) {
    S.id[i] = im[3 * (size_t)i];
    S.type[i] = im[3 * (size_t)i + 1];
    S.mark[i] = im[3 * (size_t)i + 2];
    if (S.type[i] < 1 || S.type[i] > I.ntype) return fail(c, SDEM_E_ARG, "particle type out of range in restart");
    for (int d = 0; d < 3; d++) {
      const double x = pos[3 * (size_t)i + d];
      // one rank's subdomain is the whole domain: a particle outside it
      // makes nlocal_sum /= np_InDomain (:834-837)
      if (!(x >= I.lo[d] && x < I.hi[d])) return fail(c, SDEM_E_ARG, "nlocal_sum /= np_InDomain");
      S.posr[4 * (size_t)i + d] = x;
    }
    S.posr[4 * (size_t)i + 3] = I.type_radius[S.type[i] - 1];
    const double* q = &blk[3 * (size_t)nreal3 * i];
    auto take = [&](double* dst, int cnt) {
      std::memcpy(dst, q, 3 * sizeof(double) * cnt);
      q += 3 * cnt;
    };
    take(&S.lv[3 * (size_t)ts * i], ts);
    take(&S.la[3 * (size_t)ts * i], ts);
    take(&S.th[3 * (size_t)i], 1);
    take(&S.rv[3 * (size_t)rs * i], rs);
    take(&S.ra[3 * (size_t)rs * i], rs);
  }

  // contact records
  std::unordered_map<int32_t, int> row;
  row.reserve(2 * (size_t)np);
  for (int i = 0; i < np; i++) row[S.id[i]] = i;
  std::unordered_set<int32_t> fixed(I.fix_id, I.fix_id + I.nfix), walls(I.wall_id, I.wall_id + I.nwall);
  std::vector<int32_t> kind, owner, partner;
  std::vector<double> tan, vri;
  int64_t at = 0;
  for (int i = 0; i < np; i++) {
    if (ncv[i] < 0 || at + ncv[i] > nlinks) return fail(c, SDEM_E_ARG, "link counts do not match nLinks");
    const int32_t g1 = S.id[i];
    const double* p1 = &S.posr[4 * (size_t)i];
    for (int j = 0; j < ncv[i]; j++, at++) {
      const int32_t g2 = lk[2 * at], clid = lk[2 * at + 1];
      if (clid < 1 || clid > ntan) return fail(c, SDEM_E_ARG, "CLid outside the TanDel block");
      int k = 0;
      auto it = row.find(g2);
      if (it != row.end()) {
        const double* p2 = &S.posr[4 * (size_t)it->second];
        const double dx = p1[0] - p2[0], dy = p1[1] - p2[1], dz = p1[2] - p2[2];
        const double dist = std::sqrt(dx * dx + dy * dy + dz * dz) - 1.5 * (p1[3] + p2[3]);
        if (g1 > g2 && dist <= 0.0) continue;  // the lower id's record
        bool across = false;
        const double dd[3] = {dx, dy, dz};
        for (int d = 0; d < 3; d++)
          if (I.periodic[d] && std::fabs(dd[d]) > 0.5 * (I.hi[d] - I.lo[d])) across = true;
        if (across) k = 2;
        else if (g1 < g2) k = 1;
        else continue;  // never looked up (see above)
      } else if (fixed.count(g2)) {
        k = 3;
      } else if (walls.count(g2)) {
        k = 4;
      } else {
        return fail(c, SDEM_E_ARG, "restart link to an unknown id " + std::to_string(g2));
      }
      const double* t = &td[4 * (size_t)(clid - 1)];
      kind.push_back(k);
      owner.push_back(g1);
      partner.push_back(g2);
      tan.insert(tan.end(), t, t + 3);
      vri.push_back(t[3]);
    }
  }
  if (at != nlinks) return fail(c, SDEM_E_ARG, "link counts do not match nLinks");
  if (int rc = sdem_set_state(c, np, S.id.data(), S.type.data(), S.mark.data(), S.posr.data(), S.lv.data(),
                              S.la.data(), S.th.data(), S.rv.data(), S.ra.data()))
    return rc;
  return sdem_set_contacts(c, (int64_t)kind.size(), kind.data(), owner.data(), partner.data(), tan.data(),
                           vri.data());
}

// PIO_ReadFixedCoord (Prtcl_IOAndVisu.f90:164-291), one rank
extern "C" int sdem_read_fixed_coord(sdem_ctx* c, const char* path, int32_t num_fix) {
  if (!c || !path || num_fix < 0) return SDEM_E_ARG;
  sdem_box_info I;
  if (int rc = sdem_internal_box_info(c, &I)) return rc;
  File F(path, "rb");
  if (!F.f) return fail(c, SDEM_E_ARG, std::string("cannot open file: ") + path);
  std::fseek(F.f, 0, SEEK_END);
  const long long fsize = std::ftell(F.f);
  std::fseek(F.f, 0, SEEK_SET);
  constexpr int REC = 3 * 8 + 8 + 4;  // real3 + real + int, packed (:185-189, :205)
  if (fsize != (long long)REC * num_fix) return fail(c, SDEM_E_ARG, "file byte wrong");
  std::vector<unsigned char> raw((size_t)fsize);
  if (!F.get(raw.data(), raw.size())) return fail(c, SDEM_E_ARG, "file byte wrong");
  std::vector<int32_t> id(num_fix), type(num_fix);
  std::vector<double> posr(4 * (size_t)num_fix);
  for (int k = 0; k < num_fix; k++) {
    const unsigned char* r = &raw[(size_t)REC * k];
    double x[4];
    int32_t t;
    std::memcpy(x, r, 32);
    std::memcpy(&t, r + 32, 4);
    if (t < 1 || t > I.ntype) return fail(c, SDEM_E_ARG, "fixed sphere type out of range");
    const double radius = I.type_radius[t - 1];
    if (std::fabs(2.0 * radius / x[3] - 1.0) > 1.0e-6) return fail(c, SDEM_E_ARG, "Diameter not coordinate");
    for (int d = 0; d < 3; d++)
      if (!(x[d] >= I.lo[d] && x[d] < I.hi[d])) return fail(c, SDEM_E_ARG, "nfix_sum/= numPrtclFix");
    id[k] = I.num_prtcl + k + 1;  // "NOTE HERE" (:249)
    type[k] = t;
    posr[4 * (size_t)k] = x[0];
    posr[4 * (size_t)k + 1] = x[1];
    posr[4 * (size_t)k + 2] = x[2];
    posr[4 * (size_t)k + 3] = radius;
  }
  return sdem_set_fixed_particles(c, num_fix, id.data(), type.data(), posr.data());
}

// ---------------------------------------------------------------------------
// visualisation dump and its XDMF index
// ---------------------------------------------------------------------------

// PIO_Dump_visu (Prtcl_IOAndVisu.f90:1193-1267)
extern "C" int sdem_write_visu(sdem_ctx* c, const char* path, uint32_t fields) {
  if (!c || !path) return SDEM_E_ARG;
  if (fields & ~VISU_ALL) return fail(c, SDEM_E_ARG, "unknown visualisation field");
  sdem_box_info I;
  if (int rc = sdem_internal_box_info(c, &I)) return rc;
  State S;
  if (int rc = read_state(c, I, S)) return rc;
  const int n = S.n;
  std::vector<double> force, torque;
  if (fields & (SDEM_VISU_CNTCTFORCE | SDEM_VISU_TORQUE)) {
    if (!I.keep_forces) return fail(c, SDEM_E_STATE, "contact forces are not kept (sdem_set_tuning keep_forces)");
    force.resize(3 * (size_t)n);
    torque.resize(3 * (size_t)n);
    std::vector<int32_t> fid(n);
    if (int rc = sdem_get_forces(c, n, fid.data(), force.data(), torque.data())) return rc;
  }
  File F(path, "wb");
  if (!F.f) return fail(c, SDEM_E_ARG, std::string("cannot open ") + path);
  bool ok = true;
  auto put3 = [&](const std::vector<double>& src, int levels) {
    std::vector<double> v(3 * (size_t)n);
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) v[3 * (size_t)i + d] = src[3 * (size_t)levels * i + d];  // level 1
    ok = ok && F.put(v.data(), v.size() * 8);
  };
  {
    std::vector<double> p(3 * (size_t)n);
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) p[3 * (size_t)i + d] = S.posr[4 * (size_t)i + d];
    ok = ok && F.put(p.data(), p.size() * 8);
  }
  if (fields & SDEM_VISU_ID) ok = ok && F.put(S.id.data(), 4 * (size_t)n);
  if (fields & SDEM_VISU_DIAMETER) {
    std::vector<double> dia(n);
    for (int i = 0; i < n; i++) dia[i] = 2.0 * S.posr[4 * (size_t)i + 3];
    ok = ok && F.put(dia.data(), 8 * (size_t)n);
  }
  if (fields & SDEM_VISU_TYPE) ok = ok && F.put(S.type.data(), 4 * (size_t)n);
  if (fields & SDEM_VISU_USRMARK) ok = ok && F.put(S.mark.data(), 4 * (size_t)n);
  if (fields & SDEM_VISU_LINVEL) put3(S.lv, S.ts);
  if (fields & SDEM_VISU_LINACC) put3(S.la, S.ts);
  if (fields & SDEM_VISU_THETA) put3(S.th, 1);
  if (fields & SDEM_VISU_ROTVEL) put3(S.rv, S.rs);
  if (fields & SDEM_VISU_ROTACC) put3(S.ra, S.rs);
  if (fields & SDEM_VISU_CNTCTFORCE) put3(force, 1);
  if (fields & SDEM_VISU_TORQUE) put3(torque, 1);
  if (!ok) return fail(c, SDEM_E_ARG, std::string("write failed: ") + path);
  return SDEM_OK;
}

// PIO_Init_visu's header and time list (Prtcl_IOAndVisu.f90:89-112)
extern "C" int sdem_xdmf_begin(const char* xmf_path, int32_t ifirst, int32_t ilast, int32_t save_visu) {
  if (!xmf_path || save_visu < 1 || ilast < ifirst - 1) return SDEM_E_ARG;
  File F(xmf_path, "w");
  if (!F.f) return SDEM_E_ARG;
  const int nflds = (ilast - ifirst + 1) / save_visu + 1;
  std::string s;
  s += "<?xml version=\"1.0\" ?>\n";
  s += "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n";
  s += "<Xdmf xmlns:xi=\"http://www.w3.org/2001/XInclude\" Version=\"2.0\">\n";
  s += "<Domain>\n";
  s += sp(4) + "<Grid Name=\"TimeSeries\" GridType=\"Collection\" CollectionType=\"Temporal\">\n";
  s += sp(8) + "<Time TimeType=\"List\">\n";
  s += sp(12) + "<DataItem Format=\"XML\" NumberType=\"Int\" Dimensions=\"" + fI(nflds, 6) + "\">\n";
  // the times are written without advancing, so the closing tag ends the
  // same line
  s += sp(12);
  for (int f = 1; f <= nflds; f++) s += fI((long long)(f - 1) * save_visu + ifirst - 1, 9);
  s += sp(12) + "</DataItem>\n";
  s += sp(8) + "</Time>\n";
  return F.put(s.data(), s.size()) ? SDEM_OK : SDEM_E_ARG;
}

// PIO_Write_XDMF + Write_XDMF_One (Prtcl_IOAndVisu.f90:1048-1162)
extern "C" int sdem_xdmf_grid(const char* xmf_path, const char* data_base, int32_t itime, int32_t np,
                              uint32_t fields) {
  if (!xmf_path || !data_base || np < 0 || (fields & ~VISU_ALL)) return SDEM_E_ARG;
  File F(xmf_path, "a");
  if (!F.f) return SDEM_E_ARG;
  const std::string file = std::string(data_base) + fI(itime, 10, true);
  long long disp = 0;
  std::string s;
  s += sp(8) + "<Grid Name=\"T" + fI(itime, 10, true) + "\" GridType=\"Uniform\">\n";
  s += sp(12) + "<Topology TopologyType=\"Polyvertex\" NodesPerElement=\"" + fI(np, 9) + "\"/>\n";
  s += sp(12) + "<Geometry GeometryType=\"XYZ\">\n";
  s += sp(16) + "<DataItem Format=\"Binary\" DataType=\"Float\" Precision=\"" + fI(8, 1) +
       "\" Endian=\"Native\" Dimensions=\"" + fI(3, 2) + fI(np, 9) + "\" Seek=\"" + fI(disp, 15) + "\">\n";
  disp += (long long)np * 3 * 8;
  s += sp(20) + file + "\n";
  s += sp(16) + "</DataItem>\n";
  s += sp(12) + "</Geometry>\n";
  for (const VisuField& v : VISU_FIELDS) {
    if (!(fields & v.bit)) continue;
    s += sp(12) + "<Attribute Type=\"" + v.attr + "\" Center=\"Node\" Name=\"" + v.name + "\">\n";
    s += sp(16) + "<DataItem Format=\"Binary\" DataType=\"" + v.dtype + "\" Precision=\"" + fI(v.prec, 1) +
         "\" Endian=\"Native\" Dimensions=\"" + fI(v.dims, 2) + fI(np, 9) + "\" Seek=\"" + fI(disp, 15) + "\">\n";
    disp += (long long)np * v.dims * v.prec;
    s += sp(20) + file + "\n";
    s += sp(16) + "</DataItem>\n";
    s += sp(12) + "</Attribute>\n";
  }
  s += "        </Grid>\n";
  return F.put(s.data(), s.size()) ? SDEM_OK : SDEM_E_ARG;
}

// PIO_Final_visu's tail (Prtcl_IOAndVisu.f90:1183-1185)
extern "C" int sdem_xdmf_end(const char* xmf_path) {
  if (!xmf_path) return SDEM_E_ARG;
  File F(xmf_path, "a");
  if (!F.f) return SDEM_E_ARG;
  const std::string s = "    </Grid>\n</Domain>\n</Xdmf>\n";
  return F.put(s.data(), s.size()) ? SDEM_OK : SDEM_E_ARG;
}

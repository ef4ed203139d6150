// hg_snapshot.hip -- snapshot and resume of the evolving device state.
//
// The reference's CLIs serialise their host state for a bit-identical resume
// (mpmgrains2d_cli.cpp:154-247); the DEM contact history and the hybrid's
// homogenisation accumulators are not part of what it can save.  Here a
// snapshot is a host blob:
//
//   header (magic "HGSNAP02", kind, precision, two counts, payload bytes; format 02
//   stores FP32 points' J and b_bar diagonal as their deviations from 1)
//   payload: the arrays listed per kind below, back to back.
//
//  * MPM: every per-point array of the current buffer in its physical order
//    (positions, velocities, b_bar, J, mass, volume, half width when per
//    point, ids, flags when present).  Restoring copies them into the current
//    buffer of a handle made from the same descriptors and re-sorts; P2G's
//    sums depend only on the point set (cells in order, ids within a cell),
//    so the resumed run is bitwise the snapshotted one.
//  * DEM: positions, angles and velocities in body order, and the contact
//    history of the last step (owned slots: partner keys, springs, counts;
//    plane slots: flags and springs).  Mass weights are inputs of the handle
//    (hg_dem_set_weights), not state, and are not included.
//  * hybrid: the MPM and DEM blobs, then the homogenisation flags and, when
//    on, its accumulators.
#include "hg_hybrid_handle.h"

#include <cstring>

#define HG_TRY(expr)              \
  do {                            \
    hg_status s_ = (expr);        \
    if (s_ != HG_OK) return s_;   \
  } while (0)

namespace {

constexpr char kMagic[8] = {'H', 'G', 'S', 'N', 'A', 'P', '0', '2'};
enum : uint32_t { kKindMpm = 1, kKindDem = 2, kKindHybrid = 3 };

struct SnapHeader {
  char magic[8];
  uint32_t kind, precision;
  uint64_t count0, count1;  // points and scenes / bodies and scenes / -
  uint64_t payload;         // bytes after the header
};

struct Sec {
  void *dev;
  size_t bytes;
};

uint64_t sec_bytes(const std::vector<Sec> &s) {
  uint64_t b = 0;
  for (const Sec &x : s) b += x.bytes;
  return b;
}

hg_status pack(hg_ctx *ctx, const SnapHeader &hd, const std::vector<Sec> &secs, void *buf,
               uint64_t *bytes) {
  const uint64_t need = sizeof(SnapHeader) + sec_bytes(secs);
  if (!buf || *bytes < need) {
    const bool query = !buf;
    *bytes = need;
    return query ? HG_OK : hg_fail(ctx, HG_ERR_INVALID, "snapshot buffer too small");
  }
  char *p = static_cast<char *>(buf);
  std::memcpy(p, &hd, sizeof(hd));
  p += sizeof(hd);
  for (const Sec &s : secs) {
    if (s.bytes) HG_HIP(ctx, hipMemcpy(p, s.dev, s.bytes, hipMemcpyDeviceToHost));
    p += s.bytes;
  }
  *bytes = need;
  return HG_OK;
}

hg_status check_header(hg_ctx *ctx, const void *buf, uint64_t bytes, const SnapHeader &want,
                       const std::vector<Sec> &secs) {
  if (!buf || bytes < sizeof(SnapHeader)) return hg_fail(ctx, HG_ERR_INVALID, "snapshot too short");
  SnapHeader hd;
  std::memcpy(&hd, buf, sizeof(hd));
  if (std::memcmp(hd.magic, kMagic, 8) != 0)
    return hg_fail(ctx, HG_ERR_INVALID, "not a snapshot of this engine");
  if (hd.kind != want.kind) return hg_fail(ctx, HG_ERR_INVALID, "snapshot of another kind of handle");
  if (hd.precision != want.precision)
    return hg_fail(ctx, HG_ERR_INVALID, "snapshot precision differs from the handle's");
  if (hd.count0 != want.count0 || hd.count1 != want.count1)
    return hg_fail(ctx, HG_ERR_INVALID, "snapshot sizes differ from the handle's scenes");
  if (hd.payload != sec_bytes(secs) || bytes != sizeof(SnapHeader) + hd.payload)
    return hg_fail(ctx, HG_ERR_INVALID, "snapshot length does not match the handle");
  return HG_OK;
}

hg_status unpack(hg_ctx *ctx, const void *buf, const std::vector<Sec> &secs) {
  const char *p = static_cast<const char *>(buf) + sizeof(SnapHeader);
  for (const Sec &s : secs) {
    if (s.bytes) HG_HIP(ctx, hipMemcpy(s.dev, p, s.bytes, hipMemcpyHostToDevice));
    p += s.bytes;
  }
  return HG_OK;
}

std::vector<Sec> mpm_secs(hg_mpm *h) {
  const hgmpm::PtsRaw &p = h->buf[h->cur];
  const size_t n = h->npoints, T = h->tsize();
  std::vector<Sec> s = {{p.x, n * 8},  {p.y, n * 8},   {p.vx, n * T}, {p.vy, n * T},
                        {p.b00, n * T}, {p.b01, n * T}, {p.b11, n * T}, {p.J, n * T},
                        {p.id, n * 4}};
  if (h->has_mv) {
    s.push_back({p.m, n * T});
    s.push_back({p.V, n * T});
  }
  if (h->has_hl) s.push_back({p.hl, n * T});
  if (h->has_flags) s.push_back({p.flags, n});
  return s;
}

SnapHeader mpm_header(const hg_mpm *h, uint64_t payload) {
  SnapHeader hd;
  std::memcpy(hd.magic, kMagic, 8);
  hd.kind = kKindMpm;
  hd.precision = uint32_t(h->precision);
  hd.count0 = h->npoints;
  hd.count1 = h->nscenes;
  hd.payload = payload;
  return hd;
}

std::vector<Sec> dem_secs(hg_dem *h) {
  const size_t B = h->nbodies, T = h->tsize();
  const size_t O = size_t(h->cap), P = size_t(hgdem::kMaxPlanes);
  const hgdem::SlotsRaw &sl = h->slots[h->cur];
  return {{h->b.x, B * 8},       {h->b.y, B * 8},          {h->b.th, B * 8},
          {h->b.vx, B * T},      {h->b.vy, B * T},         {h->b.om, B * T},
          {sl.okey, B * O * 4},  {sl.ods, B * O * 2 * T},  {sl.ocnt, B},
          {sl.pact, B * P},      {sl.pds, B * P * 2 * T}};
}

SnapHeader dem_header(const hg_dem *h, uint64_t payload) {
  SnapHeader hd;
  std::memcpy(hd.magic, kMagic, 8);
  hd.kind = kKindDem;
  hd.precision = uint32_t(h->precision);
  hd.count0 = h->nbodies;
  hd.count1 = h->nscenes;
  hd.payload = payload;
  return hd;
}

}  // namespace

extern "C" hg_status hg_mpm_snapshot(hg_mpm *h, void *buf, uint64_t *bytes) {
  if (!h || !bytes) return HG_ERR_INVALID;
  if (h->phase_next != 0) return hg_fail(h->ctx, HG_ERR_INVALID, "a phased step is in progress");
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  const std::vector<Sec> s = mpm_secs(h);
  return pack(h->ctx, mpm_header(h, sec_bytes(s)), s, buf, bytes);
}

extern "C" hg_status hg_mpm_restore(hg_mpm *h, const void *buf, uint64_t bytes) {
  if (!h) return HG_ERR_INVALID;
  if (h->phase_next != 0) return hg_fail(h->ctx, HG_ERR_INVALID, "a phased step is in progress");
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  const std::vector<Sec> s = mpm_secs(h);
  HG_TRY(check_header(h->ctx, buf, bytes, mpm_header(h, sec_bytes(s)), s));
  HG_TRY(unpack(h->ctx, buf, s));
  HG_HIP(h->ctx, hipMemset(h->err, 0, h->nscenes * 4));
  return hg_mpm_rebin(h);
}

extern "C" hg_status hg_dem_snapshot(hg_dem *h, void *buf, uint64_t *bytes) {
  if (!h || !bytes) return HG_ERR_INVALID;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  const std::vector<Sec> s = dem_secs(h);
  return pack(h->ctx, dem_header(h, sec_bytes(s)), s, buf, bytes);
}

extern "C" hg_status hg_dem_restore(hg_dem *h, const void *buf, uint64_t bytes) {
  if (!h) return HG_ERR_INVALID;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  const std::vector<Sec> s = dem_secs(h);
  HG_TRY(check_header(h->ctx, buf, bytes, dem_header(h, sec_bytes(s)), s));
  HG_TRY(unpack(h->ctx, buf, s));
  HG_HIP(h->ctx, hipMemset(h->err, 0, h->nscenes * 4));
  hg_dem_launch_rebin(h);
  HG_HIP(h->ctx, hipGetLastError());
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  return HG_OK;
}

// hybrid: header, u64 MPM blob length, MPM blob, u64 DEM blob length, DEM
// blob, u32 homogenisation flags, accumulators when on.
static std::vector<Sec> homog_secs(hg_hybrid *h) {
  if (!h->hg.stress) return {};
  const size_t G = h->mpm->nnode;
  return {{h->hg.stress, G * 24}, {h->hg.vel, G * 16}};
}

extern "C" hg_status hg_hybrid_snapshot(hg_hybrid *h, void *buf, uint64_t *bytes) {
  if (!h || !bytes) return HG_ERR_INVALID;
  uint64_t bm = 0, bd = 0;
  HG_TRY(hg_mpm_snapshot(h->mpm, nullptr, &bm));
  HG_TRY(hg_dem_snapshot(h->dem, nullptr, &bd));
  const std::vector<Sec> hs = homog_secs(h);
  const uint64_t need = sizeof(SnapHeader) + 8 + bm + 8 + bd + 4 + sec_bytes(hs);
  if (!buf || *bytes < need) {
    const bool query = !buf;
    *bytes = need;
    return query ? HG_OK : hg_fail(h->ctx, HG_ERR_INVALID, "snapshot buffer too small");
  }
  SnapHeader hd;
  std::memcpy(hd.magic, kMagic, 8);
  hd.kind = kKindHybrid;
  hd.precision = uint32_t(h->mpm->precision);
  hd.count0 = h->mpm->npoints;
  hd.count1 = h->dem->nbodies;
  hd.payload = need - sizeof(SnapHeader);
  char *p = static_cast<char *>(buf);
  std::memcpy(p, &hd, sizeof(hd));
  p += sizeof(hd);
  std::memcpy(p, &bm, 8);
  p += 8;
  HG_TRY(hg_mpm_snapshot(h->mpm, p, &bm));
  p += bm;
  std::memcpy(p, &bd, 8);
  p += 8;
  HG_TRY(hg_dem_snapshot(h->dem, p, &bd));
  p += bd;
  const uint32_t flags = (h->homog_stress ? 1u : 0u) | (h->homog_vel ? 2u : 0u) |
                         (h->hg.stress ? 4u : 0u);
  std::memcpy(p, &flags, 4);
  p += 4;
  HG_HIP(h->ctx, hipStreamSynchronize(h->stream));
  for (const Sec &s : hs) {
    HG_HIP(h->ctx, hipMemcpy(p, s.dev, s.bytes, hipMemcpyDeviceToHost));
    p += s.bytes;
  }
  *bytes = need;
  return HG_OK;
}

extern "C" hg_status hg_hybrid_restore(hg_hybrid *h, const void *buf, uint64_t bytes) {
  if (!h) return HG_ERR_INVALID;
  if (!buf || bytes < sizeof(SnapHeader) + 8)
    return hg_fail(h->ctx, HG_ERR_INVALID, "snapshot too short");
  SnapHeader hd;
  std::memcpy(&hd, buf, sizeof(hd));
  if (std::memcmp(hd.magic, kMagic, 8) != 0 || hd.kind != kKindHybrid)
    return hg_fail(h->ctx, HG_ERR_INVALID, "not a hybrid snapshot of this engine");
  if (bytes != sizeof(SnapHeader) + hd.payload)
    return hg_fail(h->ctx, HG_ERR_INVALID, "snapshot length does not match its header");
  const char *p = static_cast<const char *>(buf) + sizeof(hd);
  const char *end = static_cast<const char *>(buf) + bytes;
  uint64_t bm = 0, bd = 0;
  std::memcpy(&bm, p, 8);
  p += 8;
  if (bm > uint64_t(end - p)) return hg_fail(h->ctx, HG_ERR_INVALID, "corrupt hybrid snapshot");
  HG_TRY(hg_mpm_restore(h->mpm, p, bm));
  p += bm;
  if (uint64_t(end - p) < 8) return hg_fail(h->ctx, HG_ERR_INVALID, "corrupt hybrid snapshot");
  std::memcpy(&bd, p, 8);
  p += 8;
  if (bd > uint64_t(end - p)) return hg_fail(h->ctx, HG_ERR_INVALID, "corrupt hybrid snapshot");
  HG_TRY(hg_dem_restore(h->dem, p, bd));
  p += bd;
  // the next step's pre-step grain positions
  if (h->dem->nbodies) {
    HG_HIP(h->ctx, hipMemcpy(h->prex, h->dem->b.x, size_t(h->dem->nbodies) * 8,
                             hipMemcpyDeviceToDevice));
    HG_HIP(h->ctx, hipMemcpy(h->prey, h->dem->b.y, size_t(h->dem->nbodies) * 8,
                             hipMemcpyDeviceToDevice));
  }
  if (uint64_t(end - p) < 4) return hg_fail(h->ctx, HG_ERR_INVALID, "corrupt hybrid snapshot");
  uint32_t flags = 0;
  std::memcpy(&flags, p, 4);
  p += 4;
  if (flags & 4u) {
    HG_TRY(hg_hybrid_set_homogenize(h, (flags & 1u) ? 1 : 0, (flags & 2u) ? 1 : 0));
    const std::vector<Sec> hs = homog_secs(h);
    if (sec_bytes(hs) != uint64_t(end - p))
      return hg_fail(h->ctx, HG_ERR_INVALID, "corrupt hybrid snapshot");
    for (const Sec &s : hs) {
      HG_HIP(h->ctx, hipMemcpy(s.dev, p, s.bytes, hipMemcpyHostToDevice));
      p += s.bytes;
    }
  } else if (p != end) {
    return hg_fail(h->ctx, HG_ERR_INVALID, "corrupt hybrid snapshot");
  } else {
    HG_TRY(hg_hybrid_set_homogenize(h, 0, 0));
  }
  return HG_OK;
}

// Packet form of compiled programs, for one-wave workgroups: each job's steps are
// regrouped into packets of mutually independent steps that share the wave's 64 lanes
// (step k of a packet owns lanes base_k .. base_k + lanes_k - 1), so a job of s steps
// runs in about max(tree depth, Σ lanes / 64) packet rounds instead of s step rounds.
//
// Per job:
//  1. Every operand is traced to the step that wrote it (the latest earlier step whose
//     output starts at that offset) or to the leaves.
//  2. List scheduling on those true dependencies only: a step goes to the first packet
//     after all its producers' packets that still has room for its lanes (and fewer than
//     TS_PACKET_CHUNK records).  Lanes of a step: 2^out_rank, or 2^min(sum_rank, 6) if
//     larger for a general step (its summed-term offsets are spread over its lanes).
//  3. Packets are laid out in order, steps of one packet in their original order; a
//     packet that would straddle a TS_PACKET_CHUNK-record boundary is split there (the
//     kernel stages records in chunks of that many).
//  4. Intermediates get new arena offsets by first fit over their lifetimes in packets
//     (written in packet p, last read in packet q: the space is free again from packet
//     q + 1), above the leaf prefix, which is never overwritten.  No step of a packet
//     therefore writes memory that another step of the same packet reads or writes.
//
// Records keep their meaning; the packet form adds TS_STEP_PACKET to every record,
// TS_STEP_PEND to the last record of each packet, the step's first lane in pad[0], and
// in the packet's first record the 64-bit mask of first lanes in o_lpos[16..23] (little
// endian; every step of a packetized job has at most 6 out bits, so those bytes are
// otherwise unused padding).

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "tensorsat_prog.h"

namespace {

struct Live {
  int64_t off, end;
  int last;  // last packet reading it (INT32_MAX: the job result)
};

int64_t align16(int64_t x) { return (x + 15) & ~int64_t(15); }

int packetize_job(ts_job& job, ts_step* st) {
  const int n = job.num_steps;
  if (n <= 0) return TSP_OK;
  const int64_t A = job.arena_elems;
  const int64_t leaf_end = job.num_leaf_bytes;
  if (A <= 0) return TSP_EINVAL;
  std::vector<int> wr(A, -1);  // latest step whose output starts at this offset
  std::vector<int> pl(n, -1), pr(n, -1);
  std::vector<int> lanes(n);
  for (int i = 0; i < n; ++i) {
    const ts_step& s = st[i];
    if (s.out_rank > 6 || s.flags & (TS_STEP_PACKET | TS_STEP_PEND)) return TSP_EINVAL;
    const bool norhs = s.flags & TS_STEP_NORHS;
    const int64_t lsz = int64_t(1) << s.lhs_rank, rsz = int64_t(1) << s.rhs_rank;
    if (s.lhs_off + lsz > A || (!norhs && s.rhs_off + rsz > A) || s.out_off + (int64_t(1) << s.out_rank) > A)
      return TSP_EINVAL;
    pl[i] = wr[s.lhs_off];
    if (pl[i] < 0 && s.lhs_off + lsz > leaf_end) return TSP_EINVAL;  // neither written nor a leaf
    if (!norhs) {
      pr[i] = wr[s.rhs_off];
      if (pr[i] < 0 && s.rhs_off + rsz > leaf_end) return TSP_EINVAL;
    }
    wr[s.out_off] = i;
    const bool general = !(s.flags & TS_STEP_KINNER) && s.sum_rank > 0;
    const int lb = std::max<int>(s.out_rank, general ? std::min<int>(s.sum_rank, 6) : 0);
    lanes[i] = 1 << lb;
  }
  const int res = job.result_off >= 0 && job.result_off < A ? wr[job.result_off] : -1;
  if (res < 0 || st[res].out_rank != job.result_rank) return TSP_EINVAL;

  // 2. list scheduling
  std::vector<int> pk(n);
  std::vector<int> used, cnt;
  for (int i = 0; i < n; ++i) {
    int p = 0;
    if (pl[i] >= 0) p = std::max(p, pk[pl[i]] + 1);
    if (pr[i] >= 0) p = std::max(p, pk[pr[i]] + 1);
    while (p < (int)used.size() && (used[p] + lanes[i] > TS_PACKET_LANES || cnt[p] >= TS_PACKET_CHUNK)) ++p;
    if (p == (int)used.size()) {
      used.push_back(0);
      cnt.push_back(0);
    }
    used[p] += lanes[i];
    cnt[p] += 1;
    pk[i] = p;
  }
  // 3. layout with splits at chunk boundaries
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pk[a] < pk[b]; });
  std::vector<int> fp(n);  // final packet of each step
  int cur = -1, prev_pk = -1, pos = 0;
  for (int k = 0; k < n; ++k) {
    const int i = order[k];
    if (pk[i] != prev_pk || pos % TS_PACKET_CHUNK == 0) ++cur;  // new packet, or a split
    prev_pk = pk[i];
    fp[i] = cur;
    ++pos;
  }
  const int np = cur + 1;
  // 4. lifetimes and first-fit allocation above the leaves
  std::vector<int> last(n, -1);
  for (int i = 0; i < n; ++i) {
    if (pl[i] >= 0) last[pl[i]] = std::max(last[pl[i]], fp[i]);
    if (pr[i] >= 0) last[pr[i]] = std::max(last[pr[i]], fp[i]);
  }
  last[res] = INT32_MAX;
  std::vector<int64_t> noff(n, 0);
  std::vector<Live> live;  // sorted by offset
  const int64_t base = align16(leaf_end);
  int64_t top = base;
  for (int k = 0; k < n;) {
    const int P = fp[order[k]];
    live.erase(std::remove_if(live.begin(), live.end(), [&](const Live& b) { return b.last < P; }), live.end());
    for (; k < n && fp[order[k]] == P; ++k) {
      const int i = order[k];
      const int64_t sz = align16(int64_t(1) << st[i].out_rank);
      int64_t at = base;
      size_t slot = 0;
      for (; slot < live.size(); ++slot) {
        if (live[slot].off - at >= sz) break;
        at = std::max(at, live[slot].end);
      }
      noff[i] = at;
      live.insert(live.begin() + (long)slot, Live{at, at + sz, last[i] < 0 ? P : last[i]});
      top = std::max(top, at + sz);
    }
  }
  if (top > INT32_MAX) return TSP_EINVAL;
  // 5. rewrite the records in packet order
  std::vector<ts_step> out(n);
  int k = 0;
  for (int P = 0; P < np; ++P) {
    const int k0 = k;
    uint64_t mask = 0;
    int lane = 0;
    for (; k < n && fp[order[k]] == P; ++k) {
      const int i = order[k];
      ts_step r = st[i];
      r.out_off = (uint32_t)noff[i];
      if (pl[i] >= 0) r.lhs_off = (uint32_t)noff[pl[i]];
      if (pr[i] >= 0) r.rhs_off = (uint32_t)noff[pr[i]];
      if (r.flags & TS_STEP_NORHS) r.rhs_off = r.lhs_off;
      r.flags |= TS_STEP_PACKET;
      r.pad[0] = (uint8_t)lane;
      mask |= uint64_t(1) << lane;
      lane += lanes[i];
      out[k] = r;
    }
    out[k - 1].flags |= TS_STEP_PEND;
    std::memcpy(&out[k0].o_lpos[16], &mask, 8);
  }
  std::memcpy(st, out.data(), sizeof(ts_step) * (size_t)n);
  job.result_off = (int32_t)noff[res];
  job.arena_elems = top;
  return TSP_OK;
}

}  // namespace

extern "C" int tsc_packetize(ts_job* jobs, int64_t num_jobs, ts_step* steps, int64_t num_steps, int32_t threads) {
  if (num_jobs < 0 || num_steps < 0 || (num_jobs > 0 && !jobs) || (num_steps > 0 && !steps)) return TSP_EINVAL;
  for (int64_t j = 0; j < num_jobs; ++j)
    if (jobs[j].step_begin < 0 || jobs[j].num_steps < 0 || (int64_t)jobs[j].step_begin + jobs[j].num_steps > num_steps)
      return TSP_EINVAL;
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = (int)std::max<int64_t>(1, std::min<int64_t>(nt, num_jobs / 64 + 1));
  std::vector<int> rc(nt, TSP_OK);
  auto work = [&](int t) {
    for (int64_t j = t; j < num_jobs; j += nt) {
      const int r = packetize_job(jobs[j], steps + jobs[j].step_begin);
      if (r != TSP_OK) {
        rc[t] = r;
        return;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  for (int r : rc)
    if (r != TSP_OK) return r;
  return TSP_OK;
}

// ---- lane form ------------------------------------------------------------------

static_assert(sizeof(ts_lane) == 16, "one uint4 per lane record");

namespace {

uint32_t scatter(uint32_t v, int n, const uint8_t* pos) {
  uint32_t off = 0;
  for (int i = 0; i < n; ++i)
    if (((v >> i) & 1u) && pos[i] != TS_NOPOS) off |= 1u << pos[i];
  return off;
}

// Lane records of one packetized job (out == nullptr: count only).  Returns the count or
// a negative TSP_ code.
int64_t lanes_job(const ts_job& job, const ts_step* st, ts_lane* out) {
  const int n = job.num_steps;
  if (job.arena_elems > 0xFFFF) return TSP_EINVAL;
  int64_t pos = 0;
  for (int b = 0; b < n;) {
    int e = b;
    while (e < n && !(st[e].flags & TS_STEP_PEND)) ++e;
    if (e == n) return TSP_EINVAL;  // unterminated packet
    ++e;
    // lanes of the packet: past the last step's first lane + its lane count
    int nl = 0;
    for (int k = b; k < e; ++k) {
      const ts_step& s = st[k];
      if (!(s.flags & TS_STEP_PACKET) || s.out_rank > 6 || s.sum_rank > 6) return TSP_EINVAL;
      const bool general = !(s.flags & TS_STEP_KINNER) && s.sum_rank > 0;
      const int lb = std::max<int>(s.out_rank, general ? s.sum_rank : 0);
      nl = std::max(nl, (int)s.pad[0] + (1 << lb));
    }
    if (nl > TS_PACKET_LANES) return TSP_EINVAL;
    if (pos % TS_LANE_CHUNK + nl > TS_LANE_CHUNK) {  // filler to the chunk boundary
      const int skip = (int)(TS_LANE_CHUNK - pos % TS_LANE_CHUNK);
      if (out) {
        for (int i = 0; i < skip; ++i) {
          ts_lane& r = out[pos + i];
          std::memset(&r, 0, sizeof r);
          r.out = TS_LANE_NONE;
          r.op = TS_LANE_SKIP;
        }
        out[pos].count = (uint8_t)skip;
      }
      pos += skip;
    }
    if (out) {
      for (int l = 0; l < nl; ++l) {
        ts_lane& r = out[pos + l];
        std::memset(&r, 0, sizeof r);
        r.out = TS_LANE_NONE;
      }
      for (int k = b; k < e; ++k) {
        const ts_step& s = st[k];
        const bool norhs = s.flags & TS_STEP_NORHS;
        const bool general = !(s.flags & TS_STEP_KINNER) && s.sum_rank > 0;
        const int ro = s.out_rank, nc = s.sum_rank;
        const int lb = std::max<int>(ro, general ? nc : 0);
        const uint32_t rbase = norhs ? s.lhs_off : s.rhs_off;
        for (int le = 0; le < (1 << lb); ++le) {
          ts_lane& r = out[pos + s.pad[0] + le];
          const uint32_t oi = (uint32_t)le & ((1u << ro) - 1u);  // aliases past 2^ro
          r.lhs = (uint16_t)(s.lhs_off + scatter(oi, ro, s.o_lpos));
          r.rhs = (uint16_t)(norhs ? r.lhs : rbase + scatter(oi, ro, s.o_rpos));
          r.out = le < (1 << ro) ? (uint16_t)(s.out_off + (uint32_t)le) : (uint16_t)TS_LANE_NONE;
          r.op = (uint8_t)(nc | (general ? TS_LANE_GENERAL : 0) | (norhs ? TS_LANE_NORHS : 0));
          r.first = s.pad[0];
          if (general && le < (1 << nc)) {
            r.tl = (uint16_t)scatter((uint32_t)le, nc, s.c_lpos);
            r.tr = (uint16_t)(norhs ? 0 : scatter((uint32_t)le, nc, s.c_rpos));
          }
        }
      }
      out[pos].count = (uint8_t)nl;
    }
    pos += nl;
    b = e;
  }
  return pos;
}

}  // namespace

extern "C" int tsc_lanes(const ts_job* jobs, int64_t num_jobs, const ts_step* steps, int64_t num_steps, ts_lane* lanes,
                         int64_t cap, int64_t* lane_begin, int32_t threads) {
  if (num_jobs < 0 || !lane_begin || (num_jobs > 0 && (!jobs || !steps))) return TSP_EINVAL;
  for (int64_t j = 0; j < num_jobs; ++j)
    if (jobs[j].step_begin < 0 || jobs[j].num_steps < 0 || (int64_t)jobs[j].step_begin + jobs[j].num_steps > num_steps)
      return TSP_EINVAL;
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = (int)std::max<int64_t>(1, std::min<int64_t>(nt, num_jobs / 64 + 1));
  std::vector<int> rc(nt, TSP_OK);
  if (!lanes) {
    std::vector<int64_t> cnt(num_jobs);
    auto work = [&](int t) {
      for (int64_t j = t; j < num_jobs; j += nt) {
        const int64_t c = lanes_job(jobs[j], steps + jobs[j].step_begin, nullptr);
        if (c < 0) {
          rc[t] = (int)c;
          return;
        }
        cnt[j] = c;
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    for (int r : rc)
      if (r != TSP_OK) return r;
    lane_begin[0] = 0;
    for (int64_t j = 0; j < num_jobs; ++j) lane_begin[j + 1] = lane_begin[j] + cnt[j];
    return TSP_OK;
  }
  if (lane_begin[num_jobs] > cap) return TSP_EINVAL;
  auto work = [&](int t) {
    for (int64_t j = t; j < num_jobs; j += nt) {
      const int64_t c = lanes_job(jobs[j], steps + jobs[j].step_begin, lanes + lane_begin[j]);
      if (c < 0 || c != lane_begin[j + 1] - lane_begin[j]) {
        rc[t] = TSP_EINVAL;
        return;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  for (int r : rc)
    if (r != TSP_OK) return r;
  return TSP_OK;
}

// Launch recording and batched replay (see batch.h).
#include "batch.h"

#include <algorithm>
#include <cstring>
#include <string>

namespace cmx {

Recorder*& active_recorder() {
  thread_local Recorder* r = nullptr;
  return r;
}

namespace {
template <typename T> LaunchRec make_rec(int kind, const T& v, int nblocks) {
  LaunchRec r;
  r.kind = kind;
  r.nblocks = nblocks;
  r.bytes.resize(sizeof(T));
  std::memcpy(r.bytes.data(), &v, sizeof(T));
  return r;
}
}  // namespace

void record_hop(Recorder& r, const HopLaunch& h) {
  LaunchRec x = make_rec(REC_HOP, h, hop_num_blocks(h));
  // everything that selects the kernel template and the work-group size
  const int k[8] = {h.prec, h.recon, h.fast, h.prefetch, h.isign, h.epi, h.partials != nullptr, h.block};
  std::memcpy(x.key, k, sizeof(k));
  r.recs.push_back(std::move(x));
}

void record_blas(Recorder& r, int op, const BlasLaunch& b, const int prec[4]) {
  LaunchRec x = make_rec(REC_BLAS, b, blas_blocks(b.n));
  const int k[8] = {op, prec[0], prec[1], prec[2], prec[3], 0, 0, 0};
  std::memcpy(x.key, k, sizeof(k));
  r.recs.push_back(std::move(x));
}

void record_fin(Recorder& r, const FinLaunch& f) { r.recs.push_back(make_rec(REC_FIN, f, 1)); }

void BatchPlan::build(const std::vector<Recorder>& recs) {
  st_.clear();
  CMX_REQUIRE(!recs.empty(), CMX_ERR_ARG, "batch: no problems");
  const size_t ns = recs[0].recs.size();
  const int np = (int)recs.size();
  for (const Recorder& r : recs)
    CMX_REQUIRE(r.recs.size() == ns, CMX_ERR_ARG, "batch: the problems' kernel sequences differ in length");
  for (size_t j = 0; j < ns; ++j) {
    const LaunchRec& r0 = recs[0].recs[j];
    Stage s;
    s.kind = r0.kind;
    s.nprob = np;
    s.rec0 = r0.bytes;
    std::memcpy(s.key, r0.key, sizeof(s.key));
    for (int p = 0; p < np; ++p) s.nb.push_back(recs[p].recs[j].nblocks);
    if (s.kind == REC_BLAS) {
      s.opc = r0.key[0];
      for (int k = 0; k < 4; ++k) s.prec[k] = r0.key[1 + k];
    }
    std::vector<unsigned char> tab;
    tab.reserve(r0.bytes.size() * np);
    std::vector<int4> map;
    int maxnb = 0;
    for (int p = 0; p < np; ++p) {
      const LaunchRec& r = recs[p].recs[j];
      CMX_REQUIRE(r.kind == r0.kind && std::memcmp(r.key, r0.key, sizeof(r.key)) == 0 && r.bytes.size() == r0.bytes.size(),
                  CMX_ERR_ARG,
                  "batch: stage " + std::to_string(j) + " selects a different kernel for problem " + std::to_string(p) +
                      " (launch shape, precision or epilogue)");
      tab.insert(tab.end(), r.bytes.begin(), r.bytes.end());
      if (s.kind == REC_BLAS)
        for (int b = 0; b < r.nblocks; ++b) map.push_back(make_int4(p, b, r.nblocks, 0));
      maxnb = std::max(maxnb, r.nblocks);
    }
    if (s.kind == REC_HOP) {
      // Hopping stages interleave the problems in runs of 8 work-groups: the work-groups of every
      // problem over the same 8 local blocks follow each other, and global work-group g lands on XCD
      // g mod 8 = local block mod 8.  Problems on one gauge field (the sources of a propagator) then
      // read the links and clover of a block from their XCD's L2 after the first problem brought
      // them in, and the plane kernel's XCD-aware tiling keeps its meaning.  Runs shorter than 8
      // are padded with idle entries (problem -1) to keep the alignment.
      for (int j8 = 0; j8 < maxnb; j8 += 8)
        for (int p = 0; p < np; ++p) {
          const int nbp = recs[p].recs[j].nblocks;
          if (j8 >= nbp) continue;
          for (int x = 0; x < 8; ++x) {
            const int b = j8 + x;
            map.push_back(b < nbp ? make_int4(p, b, nbp, 0) : make_int4(-1, 0, 0, 0));
          }
        }
    }
    s.nblocks = s.kind == REC_FIN ? np : (int)map.size();
    s.tab.alloc(tab.size());
    CMX_HIP(hipMemcpy(s.tab.p, tab.data(), tab.size(), hipMemcpyHostToDevice));
    if (!map.empty()) {
      s.map.alloc(sizeof(int4) * map.size());
      CMX_HIP(hipMemcpy(s.map.p, map.data(), sizeof(int4) * map.size(), hipMemcpyHostToDevice));
    }
    st_.push_back(std::move(s));
  }
}

void BatchPlan::refresh(const std::vector<Recorder>& recs, hipStream_t strm) {
  const int np = (int)recs.size();
  bool same = np >= 1 && !st_.empty();
  for (int p = 0; same && p < np; ++p) same = recs[p].recs.size() == st_.size();
  for (size_t j = 0; same && j < st_.size(); ++j) {
    const Stage& s = st_[j];
    same = s.nprob == np;
    for (int p = 0; same && p < np; ++p) {
      const LaunchRec& r = recs[p].recs[j];
      same = r.kind == s.kind && r.nblocks == s.nb[p] && std::memcmp(r.key, s.key, sizeof(s.key)) == 0 &&
             r.bytes.size() == s.rec0.size();
    }
  }
  if (!same) {
    build(recs);
    return;
  }
  for (size_t j = 0; j < st_.size(); ++j) {
    Stage& s = st_[j];
    const size_t rb = s.rec0.size(), bytes = rb * np;
    if (!s.htab) {
      void* h = nullptr;
      CMX_HIP(hipHostMalloc(&h, bytes, 0));
      s.htab = std::shared_ptr<void>(h, [](void* q) { (void)hipHostFree(q); });
    }
    unsigned char* h = static_cast<unsigned char*>(s.htab.get());
    for (int p = 0; p < np; ++p) std::memcpy(h + rb * p, recs[p].recs[j].bytes.data(), rb);
    s.rec0 = recs[0].recs[j].bytes;
    CMX_HIP(hipMemcpyAsync(s.tab.p, h, bytes, hipMemcpyHostToDevice, strm));
  }
}

long BatchPlan::blocks() const {
  long n = 0;
  for (const Stage& s : st_) n += s.nblocks;
  return n;
}

void BatchPlan::run(hipStream_t strm) const {
  CMX_REQUIRE(active_recorder() == nullptr, CMX_ERR_ARG, "batch: replay while recording");
  for (const Stage& s : st_) {
    const BatchGrid bg{s.tab.p, s.map.as<int4>(), s.nblocks};
    switch (s.kind) {
      case REC_HOP: {
        HopLaunch h;
        std::memcpy(&h, s.rec0.data(), sizeof(h));
        launch_hop(h, strm, &bg);
        break;
      }
      case REC_BLAS: {
        BlasLaunch b;
        std::memcpy(&b, s.rec0.data(), sizeof(b));
        launch_blas(s.opc, b, s.prec, strm, &bg);
        break;
      }
      default:
        launch_finalize_batch(s.tab.as<FinLaunch>(), s.nprob, strm);
        break;
    }
  }
}

}  // namespace cmx

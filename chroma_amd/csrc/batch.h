// Batched solves: several problems -- solves on different lattices and operators, or several
// sources on one operator -- advance through the same sequence of kernels, each kernel launched once
// for all problems over a batched grid (BatchGrid: work-group -> (problem, work-group of that
// problem's launch)).  The per-problem launch records come from running the single-problem solver
// code with its launches recorded instead of issued, so a batched iteration performs exactly the
// arithmetic of the single-problem iteration on each problem (same kernels, same work-group
// decomposition, same reduction order).
#pragma once
#include <memory>
#include <vector>

#include "op.h"

namespace cmx {

enum RecKind { REC_HOP = 0, REC_BLAS, REC_FIN };

struct LaunchRec {
  int kind = REC_HOP;
  int key[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // kernel selection; equal across the problems of a stage
  int nblocks = 0;                        // work-groups of the problem's own launch
  std::vector<unsigned char> bytes;       // the HopLaunch, BlasLaunch or FinLaunch
};

struct Recorder {
  std::vector<LaunchRec> recs;
};

// The calling thread's recorder; while it is set, launch_hop, Op::blas and Op::finalize record
// their launches into it instead of issuing them
Recorder*& active_recorder();
void record_hop(Recorder& r, const HopLaunch& h);
void record_blas(Recorder& r, int op, const BlasLaunch& b, const int prec[4]);
void record_fin(Recorder& r, const FinLaunch& f);

struct RecordScope {
  Recorder* prev;
  explicit RecordScope(Recorder& r) : prev(active_recorder()) { active_recorder() = &r; }
  ~RecordScope() { active_recorder() = prev; }
  RecordScope(const RecordScope&) = delete;
  RecordScope& operator=(const RecordScope&) = delete;
};

// The batched stages of one recorded sequence
class BatchPlan {
 public:
  // recs[p] is problem p's recording; every recording must hold the same kernels in the same order
  void build(const std::vector<Recorder>& recs);
  // the same kernel sequence with new launch records (other fields, same kernels and work-group counts):
  // the records are copied to the device on stream s through pinned staging, the work-group maps kept;
  // anything else rebuilds the plan.  The previous run of the plan must have completed.
  void refresh(const std::vector<Recorder>& recs, hipStream_t s);
  void run(hipStream_t s) const;
  size_t stages() const { return st_.size(); }
  long blocks() const;  // work-groups over all stages

 private:
  struct Stage {
    int kind = 0, opc = 0;
    int prec[4] = {0, 0, 0, 0};
    int nprob = 0, nblocks = 0;
    std::vector<unsigned char> rec0;  // the first problem's record: selects the kernel
    DevBuf tab, map;
    int key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<int> nb;              // work-groups of each problem's own launch
    std::shared_ptr<void> htab;       // pinned staging of the records (refresh)
  };
  std::vector<Stage> st_;
};

}  // namespace cmx

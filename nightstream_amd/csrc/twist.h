// twist.h -- device-resident Twist/Shout round oracles (neo-memory/src/twist_oracle.rs).
//
// Time-lane oracles keep their sparse columns as one union support (sorted time indices) with
// the columns' values aligned to it in HBM; a round evaluates every sibling pair once and a
// fold writes the parents through a prefix-sum compaction.  Address-lane oracles replay memory
// per address bucket with a stable radix sort by bucket and a K-valued prefix scan instead of
// the reference's hash map.  Round polynomials leave the GPU as coefficient vectors.
#pragma once
#include <stdint.h>

#include <memory>
#include <vector>

#include "prover.h"

namespace ns {

struct SparseCol {
  std::vector<uint32_t> idx;  // strictly increasing
  std::vector<K2> val;        // nonzero
};

enum TsKind {
  TSK_PRODUCT = 0,
  TSK_SHOUT_VALUE,
  TSK_INDEX_ADAPTER,
  TSK_BITNESS,
  TSK_TWIST_READ,
  TSK_TWIST_WRITE,
  TSK_VAL_EVAL,
  TSK_TOTAL_INC,
  TSK_READ_ADDR,
  TSK_WRITE_ADDR,
  TSK_ADDR_LOOKUP
};

constexpr int TS_MAX_COEF = 40;  // degree bound 3 + ell_addr <= 39
constexpr int TS_GROUP = 16;     // time lanes per ragged group launch (arguments by value)

struct TsTimeArgs;
struct TsFoldArgs;
struct TsLtArgs;
struct TsSharedStream;

// Device memory of a Twist/Shout oracle, drawn from a process-wide cache of blocks in
// power-of-two size classes instead of hipMalloc / hipFree: a chunk builds and drops ~30 oracles,
// and every hipFree waits for the whole device.  A block returns to the cache only when no work
// can still use it: an oracle drains its stream before its buffers are released, and a block
// replaced by a larger one while the oracle is live waits for the device first.  Contents of a
// block are undefined when it is handed out, as with hipMalloc.
struct TsBuf {
  void* p = nullptr;
  size_t bytes = 0;
  TsBuf() = default;
  TsBuf(const TsBuf&) = delete;
  TsBuf& operator=(const TsBuf&) = delete;
  ~TsBuf();
  void ensure(size_t n);
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

class TsOracleGpu {
 public:
  TsOracleGpu();
  ~TsOracleGpu();
  TsOracleGpu(const TsOracleGpu&) = delete;
  TsOracleGpu& operator=(const TsOracleGpu&) = delete;

  // constructors (the claim, where the reference returns one, is the round-0 hypercube sum)
  void init_product(const std::vector<std::vector<K2>>& factors, size_t degree_bound);
  void init_time(TsKind kind, const std::vector<K2>& r_cycle, size_t ell_n, const SparseCol* gate, const SparseCol* a,
                 const SparseCol* b, const std::vector<SparseCol>& bits, const std::vector<K2>& r_addr,
                 const std::vector<K2>& weights, K2 init_at_r_addr, const SparseCol* term_write,
                 const SparseCol* term_inc, const std::vector<SparseCol>* term_bits);
  void init_addr(TsKind kind, const std::vector<uint32_t>& init_addrs, const std::vector<K2>& init_vals,
                 const std::vector<K2>& r_cycle, size_t ell_n, const SparseCol* has_read, const SparseCol* rv,
                 const std::vector<SparseCol>* ra_bits, const SparseCol& has_write, const SparseCol& wv_or_inc,
                 const std::vector<SparseCol>& wa_bits, const SparseCol& inc, size_t ell_addr);
  void init_lookup(const std::vector<SparseCol>& addr_bits, const SparseCol& has_lookup, const std::vector<K2>& table,
                   const std::vector<K2>& r_cycle, size_t ell_n);

  size_t num_rounds() const;
  size_t degree_bound() const { return degree_bound_; }
  void evals(const K2* xs, size_t n, K2* out);
  void fold(K2 r);
  K2 claim();

  // A round and a fold in two halves, for drivers that step many oracles together: launch_*
  // enqueues the work on this oracle's stream (the round's coefficients are copied to pinned
  // host memory); finish_* waits for it.  round_coeffs = launch_round + finish_round.
  void launch_round();
  void finish_round(std::vector<K2>& c);
  void round_coeffs(std::vector<K2>& c);
  void launch_fold(K2 r, bool defer = false);
  void finish_fold();

 private:
  friend void ts_batch_launch_round(TsOracleGpu* const* os, size_t n);
  friend void ts_batch_fold(TsOracleGpu* const* os, size_t n, K2 r);
  friend void ts_join(TsOracleGpu* const* os, size_t n);
  bool time_lane() const;
  void pre_round();
  // lt: where the Twist val_pre prefix sums of at most ONEBLK_N terms are described instead of
  // launched (*lt_described), for a group launch
  bool time_round_prepare(TsTimeArgs& A, unsigned& nb, TsLtArgs* lt, bool* lt_described);
  bool time_round_launch();
  // grp: where a fold of at most ONEBLK_N entries is described instead of launched (group path)
  bool time_fold_launch(K2 r, K2* chal_dst, TsFoldArgs* grp = nullptr);
  bool fold_begin(K2 r, bool defer, TsFoldArgs* grp);
  bool addr_round_launch(bool final);
  void addr_round(std::vector<K2>& c, bool final);
  bool product_round_launch();
  K2 time_final();

  TsKind kind_ = TSK_PRODUCT;
  hipStream_t st_ = nullptr;      // where the oracle's work goes: its own stream or its group's
  hipStream_t own_st_ = nullptr;  // from the same cache as the blocks, like the pinned mailbox
  std::shared_ptr<TsSharedStream> group_;  // the batched sum-check this oracle was last stepped in
  K2* coef_h_ = nullptr;      // pinned: TS_MAX_COEF coefficients, then the folded support size
  int round_ncoef_ = 0;      // coefficients of the launched round
  bool round_dev_ = false;   // the launched round copies them from the device
  bool round_points_ = false;  // ... as values at 0 .. round_ncoef_-1 rather than coefficients
  bool fold_pending_ = false;  // a launched time-lane fold has not reported its support size
  bool defer_fold_ = false;    // ... and the next round may read that size on the device
  bool n_on_device_ = false;   // the launched round reads the support size from nslot_
  std::vector<K2> claim_round_;     // round-0 coefficients computed for the claim, reused once
  bool claim_round_valid_ = false, round_from_claim_ = false;
  size_t degree_bound_ = 0, ell_n_ = 0, ell_addr_ = 0, bit_idx_ = 0;
  K2 prefix_eq_ = K2{1, 0};
  std::vector<K2> r_cycle_, r_addr_, chals_;
  // time lanes: columns [gate, a, b, bits...] over the union support
  size_t n_ = 0, cap_ = 1, len_ = 0, ncol_ = 0, nbits_ = 0;
  int col_a_ = -1, col_b_ = -1;
  TsBuf idx_[2], cols_[2], flags_, pos_, scan_tmp_, partials_, coef_dev_, consts_;
  TsBuf nslot_;  // the last time-lane fold's support size (u32), for a round enqueued behind it
  // pinned blocks holding the constructors' uploads until the stream synchronizes
  std::vector<std::pair<void*, size_t>> stage_;
  size_t stage_off_ = 0;
  void* stage(size_t bytes);
  void stage_release();
  void h2d_staged(void* dst, const void* src, size_t bytes);
  template <class T>
  void upload(TsBuf& d, const std::vector<T>& h, size_t min_elems = 1);
  int cur_ = 0;
  size_t scan_tmp_bytes_ = 0;
  // Twist val_pre terms (sorted by time) and their exclusive prefix sums
  K2 init_at_ = K2{0, 0};
  size_t nterms_ = 0;
  TsBuf term_t_, term_v_, term_p_, term_rank_;
  bool has_rank_ = false;
  TsBuf lt_c_[2], lt_s_[2], lt_tmp_;  // per-round LT_low contributions and their prefix sums
  std::vector<uint32_t> term_t_h_;
  std::vector<K2> term_v_h_;
  // product oracle
  size_t nf_ = 0, flen_ = 0;
  TsBuf fac_[2];
  // address lanes
  size_t nev_ = 0, ninit_tab_ = 0;
  TsBuf ev_eq_, ev_hr_, ev_rv_, ev_hw_, ev_inc_, ev_ra_, ev_wa_, ev_rw_, ev_ww_, it_[2];
  TsBuf keys_[2], vals_[2], dsort_, dscan_, sort_tmp_;
  TsBuf scratch_[9];  // constructors' temporaries, kept until the oracle is dropped
  size_t sort_tmp_bytes_ = 0;
};

// One round of the batched sum-check (neo-reductions/src/sumcheck.rs:388-407): every oracle's
// round is enqueued before the host waits for any; coeffs[i] receives oracle i's coefficients.
// All oracles must have the same number of rounds, at least one.  The time lanes of a batch
// share one stream from their first batched round on (ts_join) and share launches: ragged groups
// of up to TS_GROUP members per round kernel, per sum and per fold.  Address lanes, products and
// lookups keep their own streams and launches.
void ts_batch_round(TsOracleGpu* const* os, size_t n, std::vector<std::vector<K2>>& coeffs);
// the enqueueing half of ts_batch_round; each oracle's finish_round collects its coefficients
void ts_batch_launch_round(TsOracleGpu* const* os, size_t n);
// Fold every oracle with the shared challenge (sumcheck.rs:436-442).
void ts_batch_fold(TsOracleGpu* const* os, size_t n, K2 r);
// Puts the oracles on one stream (after their own work drained) unless they already share one.
void ts_join(TsOracleGpu* const* os, size_t n);

// Shared CPU bus helpers (neo-fold/src/memory_sidecar/cpu_bus.rs): Z is the host digit matrix
// [54][m] of the CPU witness.
void bus_decode(const int8_t* Z_host, size_t m, uint64_t* z_out);
void bus_time_cols(const int8_t* Z_host, size_t m, size_t m_in, size_t bus_base, size_t chunk, size_t bus_cols,
                   size_t steps_len, size_t ell_n, uint64_t* idx_out, uint64_t* vals_out, size_t* nnz_out);
void bus_openings(const int8_t* Z_host, size_t m, size_t m_in, size_t bus_base, size_t chunk, size_t bus_cols,
                  const K2* r, size_t ell_n, K2* y_out, K2* ys_out);

}  // namespace ns

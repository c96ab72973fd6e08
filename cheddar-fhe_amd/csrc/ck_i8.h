// Exact int8 matrix-core forms of the forward NTT and of BaseConv (see
// ck_i8.hip), and a vector-unit BaseConv used as their baseline.
#pragma once

#include "ck_bconv.h"
#include "ck_i8_arith.h"
#include "ck_internal.h"

namespace ck {

// ---- 4-step NTT on the int8 matrix cores ----------------------------------

// Per-prime constant matrices of the 4-step transform ("contexts"), each a
// 16 x 16 matrix over Z_q in MFMA operand order (see ck_i8.hip):
constexpr int kI8CtxA1 = 0;      // column pass, first 16-point stage
constexpr int kI8CtxA1Mont = 1;  // the same times R (Montgomery conversion)
constexpr int kI8CtxB1 = 2;      // + p1: column pass, second stage (twiddled)
constexpr int kI8CtxA3 = 18;     // row pass, first stage
constexpr int kI8CtxB3 = 19;     // + q1: row pass, second stage (twiddled)
constexpr int kI8Ctx = 35;

// Device tables of the int8 transform for every prime of a context, built
// on first use (ntt8_tables).
struct I8NttTabs {
  const int32_t *planes = nullptr;  // [prime][kI8Ctx][4 planes][64 lanes][4]
  const int32_t *corr = nullptr;    // [prime][2][16] offset corrections
  const uint32_t *tw = nullptr;     // [prime] psi^(c (2 brv8(p) + 1)), 16 x 16 tiles
};

// One limb of a batched int8 transform: dst = NTT(src) (x R when mont).
struct Ntt8Job {
  const uint32_t *src;
  uint32_t *dst;
  int32_t prime;
  int32_t mont;
};
static_assert(sizeof(Ntt8Job) == 24, "Ntt8Job is one 24-byte record");

// Builds (once) and returns the int8 transform tables of every prime.
int ntt8_tables(Context &c, I8NttTabs *out);
// Column pass then row pass over n jobs; col_blocks = 16-column blocks per
// workgroup of the column pass (and 16-row blocks of the row pass), a
// divisor of 16.
void launch_ntt8_cols(const Context &c, const I8NttTabs &T, const Ntt8Job *jobs, int n,
                      int col_blocks, hipStream_t st);
void launch_ntt8_rows(const Context &c, const I8NttTabs &T, const Ntt8Job *jobs, int n,
                      int row_blocks, hipStream_t st);

// ---- BaseConv on the int8 matrix cores / on the vector unit ----------------

constexpr int kI8MaxDst = 64;

// One BaseConv instance: S centred int32 source limbs -> D destination limbs.
// i8: planes holds the byte planes of the table (ck_i8.hip, bconv_i8_planes);
// valu: ftab holds D x S x {hi, lo} 16-bit halves of the same table.
struct BconvI8Inst {
  const int32_t *src;
  uint32_t *dst;
  const int64_t *planes;
  const double *ftab;
  int S, D;
  int canonical;  // 1: outputs in [0, t); 0: lazy int32 within (-t, t)
  int pad_;
  uint16_t dst_limb[kI8MaxDst];
  uint8_t dst_prime[kI8MaxDst];
};

constexpr int kMaxBconvI8Inst = 4;  // instances per kernel-argument launch
struct BconvI8Launch {
  int n;
  BconvI8Inst inst[kMaxBconvI8Inst];
};

void launch_bconv_i8_table(const Context &c, const BconvI8Inst *insts, int n, int dsts,
                           int max_s, int max_d, hipStream_t st);
// kernel-argument form (single operations: no device table)
void launch_bconv_i8(const Context &c, const BconvI8Launch &L, hipStream_t st);

// The int8 operand planes of a context BaseConv table (a ModUp group's or a
// ModDown plan's FP64 {hi, lo} table on the device, S sources x D
// destinations), built on first use from the table's words and kept in the
// context. Returns nullptr (with the error set) on failure.
const int64_t *bconv_i8_planes_for(const Context &c, const double *d_ftab, int S, int D,
                                   const uint8_t *dst_prime);
// The int8 form of a FP64-table BaseConv instance (same source, outputs
// and destinations); false on failure.
bool to_i8_inst(const Context &c, const BconvInst &in, BconvI8Inst *out);
void launch_bconv_valu_table(const Context &c, const BconvI8Inst *insts, int n, int dsts,
                             int max_s, hipStream_t st);

}  // namespace ck

// Instantiates every kernel and host pass sequence of ONE (precision, grid
// size) and exports them as an OpsTable (and, for the ensemble sizes, an
// EnsTable).  The build compiles this file once per pair, with
//   -DTG_INST_PREC=64|32 -DTG_INST_N=<N>
// and links the objects with tgdns.hip, which dispatches through the tables.
#include "tg_ops.hpp"
#include "tg_batch.hpp"

#if !defined(TG_INST_PREC) || !defined(TG_INST_N)
#error "compile with -DTG_INST_PREC=64|32 -DTG_INST_N=<grid size>"
#endif

#if TG_INST_PREC == 64
using TgReal = double;
#else
using TgReal = float;
#endif

#define TG_CAT3(a, b, c) a##b##_##c
#define TG_TABLE_NAME(pre, p, n) TG_CAT3(pre, p, n)

extern const tg::OpsTable<TgReal> TG_TABLE_NAME(tg_ops_, TG_INST_PREC, TG_INST_N);
const tg::OpsTable<TgReal> TG_TABLE_NAME(tg_ops_, TG_INST_PREC, TG_INST_N) = tg::make_table<TgReal, TG_INST_N>();

extern const tg::EnsTable<TgReal> TG_TABLE_NAME(tg_ens_, TG_INST_PREC, TG_INST_N);
const tg::EnsTable<TgReal> TG_TABLE_NAME(tg_ens_, TG_INST_PREC, TG_INST_N) = tg::make_ens_table<TgReal, TG_INST_N>();

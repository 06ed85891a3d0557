// The mixed-size ensemble launches (kb_mixed of tg_batch.hpp): one kernel
// per pass and precision holding the batched bodies of every kMixSizes
// grid size, exported as an EnsMixTable per precision.
#include "tg_batch.hpp"

extern const tg::EnsMixTable<double> tg_ens_mix_64;
extern const tg::EnsMixTable<float> tg_ens_mix_32;
const tg::EnsMixTable<double> tg_ens_mix_64 = tg::make_mix_table<double>();
const tg::EnsMixTable<float> tg_ens_mix_32 = tg::make_mix_table<float>();

// Internal interface of the many-chain executor (dqmc.hip) for the serving loop: chains
// join an ensemble, every step is one slice visit of each chain passed to it, run with one
// launch per kernel kind over those chains on the ensemble's stream.
#pragma once
#include "../../include/dqmc_hip.h"

namespace dqmc {
struct Ensemble;
int ensemble_create(int device, Ensemble** out);
void ensemble_destroy(Ensemble* e);
const char* ensemble_error(Ensemble* e);
// A built FP64 chain of the ensemble's device joins (its queued work finishes first) or
// leaves (everything queued for it has run; it gets its own stream back).
int ensemble_attach(Ensemble* e, dqmc_chain_t h);
int ensemble_detach(Ensemble* e, dqmc_chain_t h);
// One slice visit (sweep_spatial + propagate) of each chain in hs (all attached).
int ensemble_step(Ensemble* e, dqmc_chain_t* hs, int count);
long long chain_steps(dqmc_chain_t h);   // slice visits made in ensembles
int chain_slices(dqmc_chain_t h);        // M
}  // namespace dqmc

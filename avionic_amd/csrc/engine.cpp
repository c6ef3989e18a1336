// Placeholder until the batched engine lands.
#include "solver.hpp"

extern "C" {
int avionic_engine_create(const avionic_engine_config *, avionic_engine **) { return AVIONIC_E_UNSUPPORTED; }
int avionic_engine_destroy(avionic_engine *) { return AVIONIC_E_UNSUPPORTED; }
size_t avionic_request_footprint(const avionic_request_desc *) { return 0; }
int avionic_engine_submit(avionic_engine *, const avionic_request_desc *, long long *) { return AVIONIC_E_UNSUPPORTED; }
int avionic_engine_step(avionic_engine *, void *) { return AVIONIC_E_UNSUPPORTED; }
int avionic_engine_poll(avionic_engine *, long long, float *, void *) { return AVIONIC_E_UNSUPPORTED; }
int avionic_engine_cancel(avionic_engine *, long long) { return AVIONIC_E_UNSUPPORTED; }
int avionic_engine_stats(const avionic_engine *, int *, size_t *) { return AVIONIC_E_UNSUPPORTED; }
}

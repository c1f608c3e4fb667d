/* Prints the size and field offsets of the public structs of
 * include/mublis_gfx950.h, compiled by the host C compiler, for comparison
 * with the Python ctypes mirrors. */
#include <stddef.h>
#include <stdio.h>

#include "mublis_gfx950.h"

#define SIZE(t) printf("%s sizeof %zu\n", #t, sizeof(t))
#define FIELD(t, f) printf("%s.%s %zu\n", #t, #f, offsetof(t, f))

int main(void) {
  SIZE(mublis_sbatch_op_t);
  FIELD(mublis_sbatch_op_t, op);
  FIELD(mublis_sbatch_op_t, trans_a);
  FIELD(mublis_sbatch_op_t, trans_b);
  FIELD(mublis_sbatch_op_t, uplo);
  FIELD(mublis_sbatch_op_t, m);
  FIELD(mublis_sbatch_op_t, n);
  FIELD(mublis_sbatch_op_t, k);
  FIELD(mublis_sbatch_op_t, alpha);
  FIELD(mublis_sbatch_op_t, beta);
  FIELD(mublis_sbatch_op_t, a);
  FIELD(mublis_sbatch_op_t, lda);
  FIELD(mublis_sbatch_op_t, b);
  FIELD(mublis_sbatch_op_t, ldb);
  FIELD(mublis_sbatch_op_t, c);
  FIELD(mublis_sbatch_op_t, ldc);
  FIELD(mublis_sbatch_op_t, side);
  FIELD(mublis_sbatch_op_t, diag);

  SIZE(mublis_server_config_t);
  FIELD(mublis_server_config_t, hbm_budget);
  FIELD(mublis_server_config_t, streams);
  FIELD(mublis_server_config_t, slots_per_stream);
  FIELD(mublis_server_config_t, max_group);
  FIELD(mublis_server_config_t, split_flop);

  SIZE(mublis_request_t);
  FIELD(mublis_request_t, op);
  FIELD(mublis_request_t, trans_a);
  FIELD(mublis_request_t, trans_b);
  FIELD(mublis_request_t, side);
  FIELD(mublis_request_t, uplo);
  FIELD(mublis_request_t, diag);
  FIELD(mublis_request_t, m);
  FIELD(mublis_request_t, n);
  FIELD(mublis_request_t, k);
  FIELD(mublis_request_t, alpha);
  FIELD(mublis_request_t, beta);
  FIELD(mublis_request_t, a);
  FIELD(mublis_request_t, b);
  FIELD(mublis_request_t, c);

  SIZE(mublis_server_stats_t);
  FIELD(mublis_server_stats_t, submitted);
  FIELD(mublis_server_stats_t, completed);
  FIELD(mublis_server_stats_t, failed);
  FIELD(mublis_server_stats_t, launches);
  FIELD(mublis_server_stats_t, grouped_launches);
  FIELD(mublis_server_stats_t, grouped_requests);
  FIELD(mublis_server_stats_t, max_pending);
  FIELD(mublis_server_stats_t, max_inflight);
  FIELD(mublis_server_stats_t, p50_ms);
  FIELD(mublis_server_stats_t, p99_ms);
  FIELD(mublis_server_stats_t, mean_ms);
  FIELD(mublis_server_stats_t, max_ms);
  FIELD(mublis_server_stats_t, flop);
  FIELD(mublis_server_stats_t, elapsed_s);
  FIELD(mublis_server_stats_t, tflops);
  FIELD(mublis_server_stats_t, hbm_budget);
  FIELD(mublis_server_stats_t, hbm_used);
  FIELD(mublis_server_stats_t, hbm_peak);
  return 0;
}

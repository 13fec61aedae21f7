/* Prints the layout of every struct of include/fmmhip.h: one line per field,
 * "<struct> <field> <offset> <size>", then "<struct> sizeof <size> 0". Built and run by
 * tests/test_abi_layout.py to check the Julia and Python mirrors of the structs. */
#include <stddef.h>
#include <stdio.h>
#include "fmmhip.h"

#define F(S, f) printf("%s %s %zu %zu\n", #S, #f, offsetof(S, f), sizeof(((S*)0)->f))
#define Z(S) printf("%s sizeof %zu 0\n", #S, sizeof(S))

int main(void) {
    F(fmmhip_options, expansion_order); F(fmmhip_options, leaf_size_source); F(fmmhip_options, leaf_size_target);
    F(fmmhip_options, multipole_acceptance); F(fmmhip_options, list_method); F(fmmhip_options, shrink_recenter);
    F(fmmhip_options, error_kind); F(fmmhip_options, error_relative); F(fmmhip_options, error_tolerance);
    F(fmmhip_options, error_bonus); F(fmmhip_options, scalar_potential); F(fmmhip_options, gradient);
    F(fmmhip_options, hessian); F(fmmhip_options, farfield); F(fmmhip_options, nearfield);
    F(fmmhip_options, self_induced); F(fmmhip_options, precision); Z(fmmhip_options);

    F(fmmhip_stats, pmax); F(fmmhip_stats, error_success); F(fmmhip_stats, n_target_branches);
    F(fmmhip_stats, n_source_branches); F(fmmhip_stats, n_m2l); F(fmmhip_stats, n_direct); F(fmmhip_stats, n_p2p_pairs);
    F(fmmhip_stats, ms_tree); F(fmmhip_stats, ms_lists); F(fmmhip_stats, ms_p2p); F(fmmhip_stats, ms_upward);
    F(fmmhip_stats, ms_m2l); F(fmmhip_stats, ms_downward); F(fmmhip_stats, ms_total); F(fmmhip_stats, ms_estimate);
    F(fmmhip_stats, ms_gather); Z(fmmhip_stats);

    F(fmmhip_request, kind); F(fmmhip_request, kernel); F(fmmhip_request, n); F(fmmhip_request, source_buffer);
    F(fmmhip_request, options); F(fmmhip_request, influence); F(fmmhip_request, max_iterations);
    F(fmmhip_request, tolerance); F(fmmhip_request, external); Z(fmmhip_request);

    F(fmmhip_request_stats, status); F(fmmhip_request_stats, sweeps); F(fmmhip_request_stats, pmax);
    F(fmmhip_request_stats, error_success); F(fmmhip_request_stats, converged); F(fmmhip_request_stats, queue_ms);
    F(fmmhip_request_stats, run_ms); F(fmmhip_request_stats, latency_ms); F(fmmhip_request_stats, start_ms);
    F(fmmhip_request_stats, end_ms); F(fmmhip_request_stats, footprint_bytes); Z(fmmhip_request_stats);

    F(fmmhip_server_stats_t, completed); F(fmmhip_server_stats_t, queued); F(fmmhip_server_stats_t, running);
    F(fmmhip_server_stats_t, p50_ms); F(fmmhip_server_stats_t, p99_ms); F(fmmhip_server_stats_t, mean_queue_ms);
    F(fmmhip_server_stats_t, span_ms); Z(fmmhip_server_stats_t);
    return 0;
}

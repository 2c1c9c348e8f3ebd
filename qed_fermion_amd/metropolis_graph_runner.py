"""LeapfrogCmpGraphRunner: capture-once / replay of HmcSampler.leapfrog_cmp
(qed_fermion/metropolis_graph_runner.py:81-166), same API.

capture(x, p, dt, tau_mask, force_b_plaq, force_b_tau, graph_memory_pool=None, n_warmups=3) -> pool
__call__(x, p, dt, tau_mask, force_b_plaq, force_b_tau) -> (x, p, force_b_plaq, force_b_tau)

The reference's whole-trajectory MetropolisGraphRunner (9-79) calls a method
that does not exist in the reference; its role is served here by
TrajectoryEngine.capture(), which puts one full trajectory (draws, solves,
forces, leapfrog, accept) into a single hipGraph.
"""
import torch


class LeapfrogCmpGraphRunner:
    def __init__(self, hmc_sampler):
        self.hmc_sampler = hmc_sampler
        self.graph = None
        self.input_buffers = {}
        self.output_buffers = {}

    def capture(self, x, p, dt, tau_mask, force_b_plaq, force_b_tau, graph_memory_pool=None, n_warmups=3):
        ib = {"x": x, "p": p, "dt": dt, "tau_mask": tau_mask, "force_b_plaq": force_b_plaq, "force_b_tau": force_b_tau}
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(n_warmups):
                self.hmc_sampler.leapfrog_cmp(**ib)
            s.synchronize()
        torch.cuda.current_stream().wait_stream(s)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, pool=graph_memory_pool):
            out = self.hmc_sampler.leapfrog_cmp(**ib)
        self.graph = graph
        self.input_buffers = ib
        self.output_buffers = {"x": out[0], "p": out[1], "force_b_plaq": out[2], "force_b_tau": out[3]}
        return graph.pool()

    def __call__(self, x, p, dt, tau_mask, force_b_plaq, force_b_tau):
        for k, v in (("x", x), ("p", p), ("dt", dt), ("tau_mask", tau_mask), ("force_b_plaq", force_b_plaq),
                     ("force_b_tau", force_b_tau)):
            self.input_buffers[k].copy_(v)
        self.graph.replay()
        o = self.output_buffers
        return o["x"], o["p"], o["force_b_plaq"], o["force_b_tau"]

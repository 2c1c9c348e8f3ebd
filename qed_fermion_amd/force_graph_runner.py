"""ForceGraphRunner: capture-once / replay of HmcSampler.force_f_fast
(qed_fermion/force_graph_runner.py:9-95), same API.

capture(psi, boson, max_iter, graph_memory_pool=None, n_warmups=3) -> pool
__call__(psi, boson) -> (Ft, xi_t, r_err)   (static outputs: they alias on the next call)

Everything force_f_fast launches -- the hop table, the two-kernel PCG
iterations, the adjoint force kernel -- goes on torch's current stream without
host synchronisation or allocation inside the library, so the whole solve plus
force is one hipGraph.
"""
import torch


class ForceGraphRunner:
    def __init__(self, hmc_sampler):
        self.hmc_sampler = hmc_sampler
        self.graph = None
        self.input_buffers = {}
        self.output_buffers = {}

    def capture(self, psi, boson, max_iter, graph_memory_pool=None, n_warmups=3):
        input_buffers = {"psi": psi, "boson": boson}
        self.hmc_sampler.max_iter = max_iter
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(n_warmups):
                self.hmc_sampler.force_f_fast(input_buffers["psi"], input_buffers["boson"], None)
            s.synchronize()
        torch.cuda.current_stream().wait_stream(s)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, pool=graph_memory_pool):
            static_outputs = self.hmc_sampler.force_f_fast(input_buffers["psi"], input_buffers["boson"], None)
        self.graph = graph
        self.input_buffers = input_buffers
        self.output_buffers = {"Ft": static_outputs[0], "xi_t": static_outputs[1], "r_err": static_outputs[3]}
        return graph.pool()

    def __call__(self, psi, boson):
        self.input_buffers["psi"].copy_(psi)
        self.input_buffers["boson"].copy_(boson)
        self.graph.replay()
        return self.output_buffers["Ft"], self.output_buffers["xi_t"], self.output_buffers["r_err"]

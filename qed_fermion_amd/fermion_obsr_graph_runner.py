"""GEtaGraphRunner: capture-once / replay of StochaticEstimator.set_eta_G_eta_inner
(qed_fermion/fermion_obsr_graph_runner.py:110-191), same API.

capture(boson, eta, graph_memory_pool=None, n_warmups=3) -> pool;  __call__(boson, eta) -> G_eta
The observable runners of that file (T1..Tvv, spsm) wrap FFT contractions that
are outside this package.
"""
import torch


class GEtaGraphRunner:
    def __init__(self, se):
        self.se = se
        self.graph = None
        self.input_buffers = {}
        self.output_buffers = {}

    def capture(self, boson, eta, graph_memory_pool=None, n_warmups=3):
        ib = {"boson": boson, "eta": eta}
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(n_warmups):
                self.se.set_eta_G_eta_inner(ib["boson"], ib["eta"])
            s.synchronize()
        torch.cuda.current_stream().wait_stream(s)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, pool=graph_memory_pool):
            out = self.se.set_eta_G_eta_inner(ib["boson"], ib["eta"])
        self.graph = graph
        self.input_buffers = ib
        self.output_buffers = {"G_eta": out}
        return graph.pool()

    def __call__(self, boson, eta):
        self.input_buffers["boson"].copy_(boson)
        self.input_buffers["eta"].copy_(eta)
        self.graph.replay()
        return self.output_buffers["G_eta"]

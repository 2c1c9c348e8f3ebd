"""Capture-once / replay runners of qed_fermion/fermion_obsr_graph_runner.py, same API.

GEtaGraphRunner (110-191): capture(max_iter_se, graph_memory_pool=None, n_warmups=3) -> pool;
    __call__(boson, eta) -> G_eta.  capture allocates its own static inputs (boson
    [1, 2, Lx, Ly, Ltau], eta [Nrv, Ltau*Vs]) and runs the solve with the sampler's
    max_iter set to max_iter_se while it captures, restoring it afterwards (117-175).
T1/T21/T2/T4/TvvGraphRunner (365-722): capture(graph_memory_pool=None, n_warmups=3) -> pool;
    __call__(eta, G_eta, boson) -> the estimator's output (Tvv: (Tvv, Tv))
The estimator graphs hold the HIP field / k-space kernels and the rocFFT
transforms between them; inputs are copied into static buffers before each replay.
"""
import torch


class GEtaGraphRunner:
    def __init__(self, se):
        self.graph = None
        self.se = se
        self.input_buffers = {}
        self.output_buffers = {}
        self.max_iter_se = None

    def capture(self, max_iter_se, graph_memory_pool=None, n_warmups=3):
        """Capture set_eta_G_eta_inner as one graph (fermion_obsr_graph_runner.py:117-175)."""
        se = self.se
        hmc = se.hmc_sampler
        input_buffers = {
            "boson": torch.zeros((1, 2, se.Lx, se.Ly, se.Ltau), device=se.device, dtype=se.dtype),
            "eta": torch.zeros((se.Nrv, se.Ltau * se.Vs), device=se.device, dtype=se.cdtype),
        }
        max_iter_copy = hmc.max_iter
        hmc.max_iter = int(max_iter_se)
        try:
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                for _ in range(n_warmups):
                    se.set_eta_G_eta_inner(input_buffers["boson"], input_buffers["eta"])
                s.synchronize()
            torch.cuda.current_stream().wait_stream(s)
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, pool=graph_memory_pool):
                static_outputs = se.set_eta_G_eta_inner(input_buffers["boson"], input_buffers["eta"])
        finally:
            hmc.max_iter = max_iter_copy
        self.graph = graph
        self.input_buffers = input_buffers
        self.output_buffers = static_outputs
        self.max_iter_se = int(max_iter_se)
        return graph.pool()

    def __call__(self, boson, eta):
        """G_eta for boson [1, 2, Lx, Ly, Ltau] and eta [Nrv, Ltau*Vs] (fermion_obsr_graph_runner.py:177-191)."""
        self.input_buffers["boson"].copy_(boson)
        self.input_buffers["eta"].copy_(eta)
        self.graph.replay()
        return self.output_buffers


class _EstimatorGraphRunner:
    method = None

    def __init__(self, se):
        self.se = se
        self.graph = None
        self.input_buffers = {}
        self.output_buffers = None

    def capture(self, graph_memory_pool=None, n_warmups=3):
        se = self.se
        n = se.Ltau * se.Ly * se.Lx
        ib = {"eta": torch.zeros((se.Nrv, n), device=se.device, dtype=se.cdtype),
              "G_eta": torch.zeros((se.Nrv, n), device=se.device, dtype=se.cdtype),
              "boson": torch.zeros((1, 2, se.Lx, se.Ly, se.Ltau), device=se.device, dtype=se.dtype)}
        fn = getattr(se, self.method)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(n_warmups):
                fn(ib["eta"], ib["G_eta"], ib["boson"])
            s.synchronize()
        torch.cuda.current_stream().wait_stream(s)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, pool=graph_memory_pool):
            out = fn(ib["eta"], ib["G_eta"], ib["boson"])
        self.graph = graph
        self.input_buffers = ib
        self.output_buffers = out
        return graph.pool()

    def __call__(self, eta, G_eta, boson):
        self.input_buffers["eta"].copy_(eta)
        self.input_buffers["G_eta"].copy_(G_eta)
        self.input_buffers["boson"].copy_(boson)
        self.graph.replay()
        return self.output_buffers


class T1GraphRunner(_EstimatorGraphRunner):
    """fermion_obsr_graph_runner.py:365-434."""
    method = "T1"


class T21GraphRunner(_EstimatorGraphRunner):
    """fermion_obsr_graph_runner.py:436-505."""
    method = "T21"


class T2GraphRunner(_EstimatorGraphRunner):
    """fermion_obsr_graph_runner.py:507-576."""
    method = "T2"


class T4GraphRunner(_EstimatorGraphRunner):
    """fermion_obsr_graph_runner.py:578-647."""
    method = "T4"


class TvvGraphRunner(_EstimatorGraphRunner):
    """fermion_obsr_graph_runner.py:649-720."""
    method = "Tvv"

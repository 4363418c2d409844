"""Scheduling logic of the serving loop on the CPU, with a stand-in for the
device ensemble that records what it is asked to do."""
from krabmaga_amd.ensemble import RunSpec
from krabmaga_amd.serve import Server, summary
from krabmaga_amd.workloads import Request


class FakeEnsemble:
    def __init__(self, log):
        self.log = log
        self.next = 0
        self.steps = {}

    def admit(self, specs):
        ids = list(range(self.next, self.next + len(specs)))
        self.next += len(specs)
        for i in ids:
            self.steps[i] = 0
        self.log.append(("admit", ids))
        return ids

    def run(self, n):
        assert n > 0
        for k in self.steps:
            self.steps[k] += n
        self.log.append(("run", n))

    def sync(self):
        pass

    def read(self, rid):
        return {"steps": self.steps[rid]}

    def retire(self, rid):
        self.log.append(("retire", rid, self.steps.pop(rid)))

    def close(self):
        pass


class Clock:
    """Advances 1 ms per call, so arrivals are reached deterministically."""
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t


def spec(n):
    return RunSpec(width=100.0, height=100.0, agents=n)


def test_every_run_stops_at_its_own_step_count():
    reqs = [Request(0.0, spec(10), 70), Request(0.0, spec(20), 5), Request(0.002, spec(30), 130)]
    log = []
    srv = Server(10**9, max_tick=64, keep_state=True, clock=Clock(),
                 ensemble=lambda: FakeEnsemble(log), footprint_fn=lambda s: s.agents)
    done = srv.serve(reqs)
    assert sorted(c.index for c in done) == [0, 1, 2]
    retired = {e[1]: e[2] for e in log if e[0] == "retire"}
    # run ids follow admission order
    by_index = {c.index: c for c in done}
    assert retired[0] == 70 and retired[1] == 5 and retired[2] == 130
    assert all(c.state["steps"] == reqs[c.index].steps for c in done)
    assert all(e[1] <= 64 for e in log if e[0] == "run")
    s = summary(done, srv.wall)
    assert s["requests"] == 3 and s["p99_latency_s"] >= s["p50_latency_s"] > 0
    assert by_index[1].done <= by_index[0].done


def test_admission_respects_the_memory_budget_in_arrival_order():
    reqs = [Request(0.0, spec(60), 10), Request(0.0, spec(50), 10), Request(0.0, spec(10), 10)]
    log = []
    srv = Server(100, clock=Clock(), ensemble=lambda: FakeEnsemble(log),
                 footprint_fn=lambda s: s.agents)
    done = srv.serve(reqs)
    assert len(done) == 3 and srv.peak_bytes <= 100
    admits = [e[1] for e in log if e[0] == "admit"]
    # the 50-agent request waits for the 60-agent one (FIFO, no overtaking)
    assert admits[0] == [0]
    assert sum(len(a) for a in admits) == 3


def test_oversized_request_is_rejected():
    import pytest
    srv = Server(10, clock=Clock(), ensemble=lambda: FakeEnsemble([]),
                 footprint_fn=lambda s: s.agents)
    with pytest.raises(ValueError):
        srv.serve([Request(0.0, spec(11), 5)])

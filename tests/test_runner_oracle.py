"""The scene runner's frame loop and records, driven through the CPU oracle (same Python API as
the library: create, seed, step, info, debug_stats): report lines per interval, step numbering,
interval cuts at checkpoint frames, resume numbering. Checkpoint files are stand-ins here (the
oracle has no state format); the library's state and resume are checked in the GPU tests."""
import json

import pytest

import dcgrid_amd
import oracle
from dcgrid_amd import run


class OracleRunSim(oracle.OracleSim):
    """The oracle with the few calls the runner uses besides the step interface."""

    def __init__(self, kind, size, p, b):
        super().__init__(kind, size, p, b, precision=32)
        self.kind = kind
        self.cells = 0
        self.saved = []

    def synchronize(self):
        pass

    def cell_steps(self):
        return self.cells

    def step(self, n=1):
        for _ in range(n):
            super().step(1)
            self.cells += int(self.info().active_cells)

    def save_state(self, path):
        self.saved.append(path)
        open(path, "w").write("stand-in")

    def load_state(self, path):
        assert open(path).read() == "stand-in"

    def set_graph(self, enable=True):
        pass

    def close(self):
        if self._h:
            inf = self.info()
            self.final_levels, self.final_active = int(inf.levels), int(inf.active_cells)
        super().close()


@pytest.fixture
def oracle_sims(monkeypatch):
    made = []

    def factory(*a):
        s = OracleRunSim(*a)
        made.append(s)
        return s

    monkeypatch.setattr(dcgrid_amd, "Simulation", factory)
    return made


def lines(path):
    return [json.loads(x) for x in open(path)]


@pytest.mark.parametrize("config,scale", [(1, 4), (4, 8), (0, 2)])
def test_records(tmp_path, oracle_sims, config, scale):
    j = tmp_path / "r.jsonl"
    assert run.main(["--config", str(config), "--scale", str(scale), "--steps", "5", "--every", "2",
                     "--residual", "--jsonl", str(j)]) == 0
    rec = lines(j)
    assert rec[0]["event"] == "start" and rec[0]["kind"] == ("uniform" if config == 0 else "dcgrid")
    assert [r["step"] for r in rec[1:-1]] == [2, 4, 5] and [r["frames"] for r in rec[1:-1]] == [2, 2, 1]
    sim = oracle_sims[0]
    assert rec[-2]["active_cells"] == sim.final_active
    for r in rec[1:-1]:
        assert r["residual"] >= 0.0 and r["active_cell_steps_per_s"] > 0
        if config != 0:
            assert len(r["block_loads"]) == sim.final_levels and sum(r["block_loads"]) * 512 == r["active_cells"]
    end = rec[-1]
    assert end["event"] == "end" and end["steps"] == 5 and end["active_cell_steps"] == sim.cell_steps()


def test_checkpoint_cuts_and_resume_numbering(tmp_path, oracle_sims):
    ck, j = tmp_path / "ck", tmp_path / "a.jsonl"
    assert run.main(["--config", "1", "--scale", "4", "--steps", "7", "--every", "5", "--checkpoint", str(ck),
                     "--checkpoint-every", "3", "--jsonl", str(j)]) == 0
    assert [r["step"] for r in lines(j)[1:-1]] == [3, 6, 7]
    assert oracle_sims[0].saved == [str(ck)] * 3  # frames 3 and 6, and the end
    j2 = tmp_path / "b.jsonl"
    assert run.main(["--config", "1", "--scale", "4", "--resume", str(ck), "--start-step", "7", "--steps", "2",
                     "--every", "1", "--jsonl", str(j2)]) == 0
    r = lines(j2)
    assert r[0]["resumed_from"] == str(ck) and [x["step"] for x in r[1:-1]] == [8, 9]

"""Argument handling of the scene runner (no GPU)."""
import pytest

from dcgrid_amd import run


def test_defaults_and_checks():
    a = run.parse_args([])
    assert a.config == 1 and a.every == 10 and a.steps is None and a.storage == "fp32"
    a = run.parse_args(["--config", "4", "--checkpoint", "s", "--checkpoint-every", "50", "--resume", "t",
                        "--start-step", "150"])
    assert (a.config, a.checkpoint, a.checkpoint_every, a.resume, a.start_step) == (4, "s", 50, "t", 150)
    for bad in (["--every", "0"], ["--adjoint"], ["--config", "7"], ["--storage", "fp8"]):
        with pytest.raises(SystemExit):
            run.parse_args(bad)

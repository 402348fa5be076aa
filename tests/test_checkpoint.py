"""Checkpoint / restart (io/__init__.py:43-122 API) on host tensors: write
every checkevery steps, keep the previous file, restore each time level
bit for bit; a killshenfun file makes update write and exit(1)."""
import os

import numpy as np
import pytest
import torch

from shenfun_amd.io import Checkpoint


def test_checkpoint_round_trip(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    rng = np.random.default_rng(0)
    U = torch.from_numpy(rng.standard_normal((4, 6)) + 1j * rng.standard_normal((4, 6)))
    U0 = torch.from_numpy(rng.standard_normal((4, 6)) + 1j * rng.standard_normal((4, 6)))
    phi = torch.from_numpy(rng.standard_normal(5))
    chk = Checkpoint("run", checkevery=2, data={"0": {"U": [U], "phi": [phi]}, "1": {"U": [U0]}})
    chk.update(0.1, 1)  # not a multiple of 2: nothing written
    assert not os.path.exists("run.chk.npz")
    chk.update(0.2, 2)
    assert chk.attrs() == (2, 0.2)
    saved = U.clone()
    U.mul_(3.0)
    chk.update(0.4, 4)
    assert os.path.exists("run.old.chk.npz") and chk.attrs() == (4, 0.4)
    a, b, p = torch.zeros_like(U), torch.zeros_like(U0), torch.zeros_like(phi)
    chk.read(a, "U")
    chk.read(b, "U", step=1)
    chk.read(p, "phi")
    assert torch.equal(a, U) and torch.equal(b, U0) and torch.equal(p, phi)
    old = Checkpoint("run.old")
    old.read(a, "U")
    assert torch.equal(a, saved)
    with pytest.raises(KeyError):
        chk.read(a, "V")
    with pytest.raises(ValueError):
        chk.read(torch.zeros(3), "phi")


def test_checkpoint_kill_file(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    U = torch.arange(6.0)
    chk = Checkpoint("k", checkevery=100, data={"0": {"U": [U]}})
    open("killshenfun", "w").close()
    with pytest.raises(SystemExit):
        chk.update(1.5, 7)
    assert not os.path.exists("killshenfun") and chk.attrs() == (7, 1.5)
    out = torch.zeros(6)
    assert torch.equal(chk.read(out, "U"), U)


def test_checkpoint_flush_keeps_a_complete_file(tmp_path, monkeypatch):
    """A stop between writing .old and renaming the new archive leaves the
    current .chk.npz in place (complete), never no file."""
    monkeypatch.chdir(tmp_path)
    U = torch.arange(4.0)
    chk = Checkpoint("c", checkevery=1, data={"0": {"U": [U]}})
    chk.update(0.1, 1)
    real_replace = os.replace
    calls = []

    def failing_replace(src, dst):
        calls.append((src, dst))
        if dst.endswith("c.chk.npz"):
            raise OSError("stopped before the final rename")
        return real_replace(src, dst)

    U.add_(1.0)
    monkeypatch.setattr(os, "replace", failing_replace)
    with pytest.raises(OSError):
        chk.update(0.2, 2)
    monkeypatch.setattr(os, "replace", real_replace)
    assert os.path.exists("c.chk.npz") and chk.attrs() == (1, 0.1)
    assert os.path.exists("c.old.chk.npz")

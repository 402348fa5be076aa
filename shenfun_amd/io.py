"""Checkpoint / restart of resident solution state (shenfun's io.Checkpoint,
io/__init__.py:43-122).

The reference writes HDF5 through h5py; h5py is not available here, so the
same API stores a NumPy ``.npz`` archive (no pickled objects: plain arrays
only).  Spectral arrays live on the GPU; ``update`` copies them to the host
once per checkpoint, ``read`` copies back into the caller's array.
Datasets are named ``name/step`` as the reference's groups, plus the
attributes ``tstep`` and ``t``.  The previous file is kept as
``<filename>.old.chk.npz`` (the reference copies ``.chk.h5`` to
``.old.chk.h5``, io/__init__.py:75-93), and the new file replaces the current
one by a rename, so a complete checkpoint exists at every moment.
"""
import os
import shutil
import sys

import numpy as np
import torch

__all__ = ["Checkpoint"]


class Checkpoint:
    """``data = {'0': {'U': [U_hat]}, '1': {'U': [U0_hat]}, ...}``: arrays of
    several time levels (0 the current one) written every ``checkevery``
    steps by ``update(t, tstep)``; ``read(u, name, step=0)`` restores one."""

    def __init__(self, filename, checkevery=10, data=None):
        self.filename = filename
        self.checkevery = int(checkevery)
        self.data = {} if data is None else data
        self.f = None

    @property
    def path(self):
        return self.filename + ".chk.npz"

    def open(self, mode="r"):
        """Load the archive into ``self.f`` (a dict of host arrays)."""
        if mode == "w" or not os.path.exists(self.path):
            self.f = {"tstep": np.array(0), "t": np.array(0.0)}
        else:
            with np.load(self.path, allow_pickle=False) as z:
                self.f = {k: z[k] for k in z.files}
        return self.f

    def close(self):
        self.f = None

    def write(self, step, d):
        """Stage ``{name: [array, ...]}`` of time level ``step`` into ``self.f``."""
        if self.f is None:
            self.open()
        for name, val in d.items():
            for u in val:
                a = u.detach().cpu().numpy() if isinstance(u, torch.Tensor) else np.asarray(u)
                self.f[f"{name}/{int(step)}"] = np.array(a)

    def _flush(self):
        """Write the new archive to a temporary file (synced to disk), keep a
        copy of the current one as .old (it stays in place meanwhile), then
        rename the new one over it: at every moment a complete .chk.npz
        exists, the old or the new."""
        tmp = self.path + ".tmp.npz"
        with open(tmp, "wb") as fh:
            np.savez(fh, **self.f)
            fh.flush()
            os.fsync(fh.fileno())
        if os.path.exists(self.path):
            old = self.filename + ".old.chk.npz"
            otmp = old + ".tmp"
            shutil.copy2(self.path, otmp)
            os.replace(otmp, old)
        os.replace(tmp, self.path)

    def update(self, t, tstep):
        """Write every time level of ``data`` when ``tstep`` is a multiple of
        ``checkevery`` or a file ``killshenfun`` is in the working directory
        (then exit with status 1 after writing, as the reference does)."""
        kill = self.check_if_kill()
        if tstep % self.checkevery == 0 or kill:
            self.open()
            for key, val in self.data.items():
                self.write(int(key), val)
            self.f["tstep"] = np.array(int(tstep))
            self.f["t"] = np.array(float(t))
            self._flush()
            self.close()
            if kill:
                sys.exit(1)

    def read(self, u, name, **kw):
        """u[...] = the stored ``name`` at time level ``step`` (default 0)."""
        step = int(kw.get("step", 0))
        self.open()
        key = f"{name}/{step}"
        if key not in self.f:
            raise KeyError(f"{key} is not in {self.path}")
        a = self.f[key]
        self.close()
        if tuple(a.shape) != tuple(u.shape):
            raise ValueError(f"stored {key} has shape {a.shape}, the array {tuple(u.shape)}")
        if isinstance(u, torch.Tensor):
            u.copy_(torch.from_numpy(a).to(u.dtype))
        else:
            u[...] = a
        return u

    def attrs(self):
        """(tstep, t) of the last checkpoint (0, 0.0 without a file)."""
        f = self.open()
        out = int(f["tstep"]), float(f["t"])
        self.close()
        return out

    @staticmethod
    def check_if_kill():
        """A file named killshenfun in the working directory asks for a clean stop."""
        if "killshenfun" in os.listdir(os.getcwd()):
            os.remove("killshenfun")
            print("killshenfun Found! Stopping simulations cleanly by checkpointing...")
            return True
        return False

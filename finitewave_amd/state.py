"""Checkpoints in the reference's on-disk format: one ``<var>.npy`` per
``state_vars`` entry in a directory.  Mirrors ``core/state/state_saver.py``
and ``core/state/state_loader.py`` (the loader restores arrays only, not
``t`` or ``step``)."""
from __future__ import annotations

from pathlib import Path

import numpy as np


class StateSaver:
    """core/state/state_saver.py:6-91"""

    def __init__(self, path=".", time=-1):
        self.path = path
        self.passed = False
        self.model = None
        self.time = time

    def initialize(self, model):
        self.model = model
        self.passed = self.path == ""

    def due(self, model) -> bool:
        """Whether save() would write at `model`'s current time (the gate of
        state_saver.py:52-77): once, at t_max or from `time` on."""
        if self.passed:
            return False
        if self.time < 0 and model.t < model.t_max:
            return False
        return not (self.time >= 0 and model.t < self.time)

    def save(self):
        if not self.due(self.model):
            return
        if not Path(self.path).exists():
            Path(self.path).mkdir(parents=True, exist_ok=True)
        for var in self.model.state_vars:
            self._save_variable(Path(self.path).joinpath(var + ".npy"),
                                self.model.__dict__[var])
        self.passed = True

    def _save_variable(self, var_path, var):
        np.save(var_path, var)


class StateSaverCollection(StateSaver):
    """core/state/state_saver.py:94-122"""

    def __init__(self):
        super().__init__()
        self.savers = []

    def initialize(self, model):
        for saver in self.savers:
            saver.initialize(model)

    def due(self, model) -> bool:
        return any(saver.due(model) for saver in self.savers)

    def save(self):
        for saver in self.savers:
            saver.save()


class StateLoader:
    """core/state/state_loader.py:6-80"""

    def __init__(self, path=""):
        self.path = path
        self.passed = True
        self.model = None

    def initialize(self, model):
        self.model = model
        self.passed = self.path == ""
        if not Path(self.path).exists():
            raise FileNotFoundError(f"Unable to load state from {self.path}. "
                                    "Directory does not exist.")

    def load(self):
        if self.passed:
            return
        for var in self.model.state_vars:
            val = self._load_variable(Path(self.path).joinpath(var + ".npy"))
            setattr(self.model, var, val)
        self.passed = True

    def _load_variable(self, var_path):
        return np.load(var_path)

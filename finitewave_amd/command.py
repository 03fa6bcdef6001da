"""Commands: user hooks run once after the first step that ends with
``t >= command.t``.  Mirrors ``core/command/command.py`` and
``core/command/command_sequence.py``.  A command sees (and may rewrite) the
host numpy state; the engine copies it back to the device and rebuilds the
plan afterwards, so mesh, fibre or weight changes take effect."""
from __future__ import annotations


class Command:
    """core/command/command.py:4-52"""

    def __init__(self, time=None):
        self.t = time
        self.passed = False

    def execute(self, model):
        raise NotImplementedError

    def update_status(self, model):
        self.passed = model.t >= self.t
        return self.passed


class CommandSequence:
    """core/command/command_sequence.py:1-59"""

    def __init__(self):
        self.sequence = []
        self.model = None

    def initialize(self, model):
        self.model = model
        for command in self.sequence:
            command.passed = False

    def add_command(self, command):
        self.sequence.append(command)

    def remove_commands(self):
        self.sequence = []

    def execute_next(self):
        for command in self.sequence:
            if not command.passed and command.update_status(self.model):
                command.execute(self.model)

"""Learned warm starts on the economic-dispatch family (kept for existing
imports; the implementation is family-generic in :mod:`.warmstart`)."""

from .warmstart import (Labels, LearnedStart, Normaliser, label, output_blocks, predict,  # noqa: F401
                        solve_family, solve_learned, step_label, targets_of, train, warm_starts)

"""A Junction whose runs go through the CPU oracle instead of the GPU --
test infrastructure for checking the oracle against the reference's golden
vectors with the same user-level scripts."""
from __future__ import annotations

import oracle
from cmtj_amd import Junction


class OracleJunction(Junction):
    def _execute(self, j, L, consts, tables, pid_out=None):
        return oracle.run(j, L, consts, tables, pid_out=pid_out)


from cmtj_amd.stack import ParallelStack, SeriesStack  # noqa: E402


class OracleSeriesStack(SeriesStack):
    def _execute(self, stack, js, layers, consts, tables=None):
        return oracle.run_stack(stack, js, layers, consts, tables)


class OracleParallelStack(ParallelStack):
    def _execute(self, stack, js, layers, consts, tables=None):
        return oracle.run_stack(stack, js, layers, consts, tables)


from cmtj_amd.reservoir import GroupInteraction  # noqa: E402


class OracleGroupInteraction(GroupInteraction):
    def _execute(self, g, js, layers, consts, pair_terms=None):
        return oracle.run_group(g, js, layers, consts, pair_terms)

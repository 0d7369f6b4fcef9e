"""kwant_amd: kwant's KPM moment engine rebuilt for AMD MI355X (gfx950).

``kwant_amd.kpm`` is a drop-in for ``kwant.kpm`` and ``kwant_amd.operator.Density``
for ``kwant.operator.Density`` (site-resolved densities evaluated inside the
recursion kernel); the numerical work runs in
the HIP kernels of ``libkpm_amd.so`` (sources in ``kwant_amd/csrc``).
"""
from . import kpm  # noqa: F401
from . import operator  # noqa: F401
from . import assembly  # noqa: F401

__all__ = ['kpm', 'operator', 'assembly']

"""Stage limiters of solve_lwfr.

PositivityPreservingLimiterZhangShu(thresholds, variables) as TrixiLW calls it
(src/callbacks/callbacks_stage/positivity_zhang_shu.jl:10-41,
positivity_zhang_shu_dg2d.jl:8-45): limiter!(u_ode, integrator, semi, t) pulls
every element whose minimum density (then pressure) is below its threshold
toward the element mean.  It runs on the device (tlw_apply_positivity_limiter);
inside the device-resident loop it is applied after every update.
"""
from __future__ import annotations

import math

density = "density"
pressure = "pressure"


class PositivityPreservingLimiterZhangShu:
    def __init__(self, thresholds=(5.0e-6, 5.0e-6), variables=(density, pressure)):
        thresholds, variables = tuple(thresholds), tuple(variables)
        if len(thresholds) != len(variables) or not variables:
            raise ValueError("one threshold per variable")
        if any(v not in (density, pressure) for v in variables) or len(set(variables)) != len(variables):
            raise NotImplementedError("the device limiter handles density and pressure")
        if variables not in ((density,), (pressure,), (density, pressure)):
            raise NotImplementedError("density is limited before pressure, as in (density, pressure)")
        thr = dict(zip(variables, thresholds))
        # a variable that is not listed is never limited (its minimum is never below -inf)
        self.device_thresholds = (float(thr.get(density, -math.inf)), float(thr.get(pressure, -math.inf)))
        self.thresholds, self.variables = thresholds, variables

    def __call__(self, u_ode, integrator, semi, t):
        semi.apply_positivity_limiter(u_ode, self.device_thresholds)

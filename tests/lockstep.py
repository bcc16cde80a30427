"""Stage-by-stage comparison of the GPU pipeline with the oracle.

Process::DoIt (Process.cpp:34-92, 238-319) is driven on both
implementations at once.  Before every stage the oracle is re-seeded with the
GPU's state (parameters and tilts), so that each stage is compared from
identical inputs:

- coarse sweeps: every pair's shift (all pairs are measured on one corrected
  stack, so no difference can compound within a sweep);
- axis searches: the whole candidate score curve, and the decision -- the
  same candidate, or a tie, where the oracle's own scores of the two picks
  differ by no more than the two curves differ anywhere;
- the offset scan: the 31-step curve and the best step;
- projection-matching sweeps: each target is measured from the references
  the previous targets left.  A sweep re-corrects every target it measures,
  and those images are the next targets' references, so a difference in one
  measurement moves every later one; the oracle therefore applies the GPU's
  measured shift of each target (orc_projalign_iter_forced) and reports its
  own measurement beside it.  The per-target difference is then the
  difference of two measurements of the same images;
- the final correction of the aligned stack.

The GPU's own sweep errors drive ProjAlignOnce's breaks, as in mf_run.
"""
from dataclasses import dataclass, field

import numpy as np

from conftest import flush_print, rel_l2

SHIFT_TOL = 0.05   # px, per result
AXIS_TOL = 0.05    # degrees


@dataclass
class Report:
    coarse_shift: list = field(default_factory=list)   # max |d| per sweep
    axis: list = field(default_factory=list)   # (range, gpu, oracle, tie)
    axis_curve: list = field(default_factory=list)  # max |d score| / peak
    offset: tuple = None
    offset_curve: float = 0.0
    pa_meas: list = field(default_factory=list)  # max |d| per sweep
    pa_params: list = field(default_factory=list)
    stack_rel: float = 0.0
    aligned_rel: float = 0.0

    @property
    def ties(self):
        return sum(1 for a in self.axis if a[3])

    def summary(self):
        return (f"coarse<= {max(self.coarse_shift, default=0):.2e} px; "
                f"axis decisions {len(self.axis)} (ties {self.ties}), "
                f"max curve diff {max(self.axis_curve, default=0):.2e}; "
                f"offset {self.offset}, curve {self.offset_curve:.2e}; "
                f"PA per-target <= {max(self.pa_meas, default=0):.2e} px "
                f"over {len(self.pa_meas)} sweeps; aligned {self.aligned_rel:.1e}")


def _grid(axis0, arange, num=100):
    step = np.float32(arange) / np.float32(num - 1)
    g = np.empty(num, np.float32)
    g[0] = np.float32(axis0) - step * np.float32(num) / np.float32(2)
    for i in range(1, num):
        g[i] = g[i - 1] + step
    return g


def _nearest(grid, axis):
    """Index of the candidate a (possibly wrapped by +-180) axis came from."""
    d = np.abs(((grid - np.float32(axis)) + 90.0) % 180.0 - 90.0)
    return int(np.argmin(d))


class Lockstep:
    def __init__(self, mf, orc, ctx, stack, tilts, log=flush_print):
        self.orc = orc
        self.G = mf.Series(ctx, stack, tilts)
        self.O = orc.Series(stack, tilts)
        self.nz = len(tilts)
        self.r = Report()
        self.log = log

    def close(self):
        self.G.close()

    def sync(self):
        sx, sy, rot, off = self.G.params()
        self.O.set_params(sx, sy, rot, off)
        self.O.set_tilts(self.G.tilts())
        return sx, sy, rot, off

    def _shift_diff(self):
        gsx, gsy, _, _ = self.G.params()
        osx, osy, _, _ = self.O.params()
        return float(max(np.abs(gsx - osx).max(), np.abs(gsy - osy).max()))

    # ------------------------------------------------------------ stages
    def preprocess(self):
        self.G.preprocess()
        self.O.preprocess()
        self.r.stack_rel = rel_l2(self.G.stack(), self.O.stack())
        assert self.r.stack_rel <= 1e-6, self.r.stack_rel
        np.testing.assert_array_equal(self.G.tilts(), self.O.tilts())

    def coarse_setup(self):
        self.G.coarse_setup()
        self.O.coarse_setup()

    def coarse_iter(self):
        self.sync()
        eg = self.G.coarse_iter()
        eo = self.O.coarse_iter()
        d = self._shift_diff()
        self.r.coarse_shift.append(d)
        self.log(f"  coarse: max shift diff {d:.2e} px, maxerr {eg:.3f}/{eo:.3f}")
        assert d <= SHIFT_TOL, d
        assert abs(eg - eo) <= SHIFT_TOL

    def tilt_axis(self, arange):
        sx, sy, rot, _ = self.sync()
        tilts = self.O.tilts()
        z0 = int(np.argmin(np.abs(tilts)))
        ag = self.G.tilt_axis(arange)
        sg = self.G.axis_scores()
        ao, so, _ = self.orc.tilt_axis(self.O.stack(), tilts,
                                       sx.astype(np.float32),
                                       sy.astype(np.float32), rot, arange)
        curve = float(np.abs(sg - so).max())
        peak = float(np.abs(so).max())
        self.r.axis_curve.append(curve / peak)
        assert curve <= 1e-5 * peak, (arange, curve, peak)
        grid = _grid(rot[z0], arange)
        kg, ko = _nearest(grid, ag), _nearest(grid, ao)
        tie = kg != ko
        if tie:
            gap = abs(float(so[ko]) - float(so[kg]))
            assert gap <= curve, ("axis", arange, ag, ao, gap, curve)
        else:
            assert abs(ag - ao) <= AXIS_TOL, (ag, ao)
        self.r.axis.append((arange, ag, ao, tie))
        self.log(f"  axis range {arange:.2f}: {ag:.4f}/{ao:.4f} "
                 f"curve diff {curve / peak:.1e}{' TIE' if tie else ''}")
        return ag

    def find_offset(self):
        self.sync()
        bg = self.G.find_offset()
        bo = self.O.find_offset()
        cg = self.G.offset_scores()
        # The oracle's curve for the same state (its scan alone).
        sx, sy, rot, _ = self.G.params()
        _, co = self.orc.find_offset(self.O.stack(), self.O.tilts(),
                                     sx.astype(np.float32),
                                     sy.astype(np.float32), rot)
        self.r.offset = (bg, bo)
        self.r.offset_curve = float(np.abs(cg - co).max() / np.abs(co).max())
        self.log(f"  offset {bg}/{bo}, curve diff {self.r.offset_curve:.1e}")
        assert bg == bo
        assert self.r.offset_curve <= 1e-5

    def apply_offset(self):
        self.G.apply_offset()
        self.O.apply_offset()
        np.testing.assert_array_equal(self.G.tilts(), self.O.tilts())

    def projalign_setup(self, align_z, nproj=10, zshift=0.0):
        sx, sy, rot, off = self.G.params()
        z = np.zeros(self.nz)
        self.G.set_params(z, z, rot, off)   # ResetShift (quirk Q16)
        self.sync()
        self.G.projalign_setup(zshift, align_z, nproj)
        self.O.projalign_setup(zshift, align_z, nproj)

    def projalign_iter(self):
        self.sync()
        eg = self.G.projalign_iter()
        meas, _ = self.G.pa_measurements()
        _, own = self.O.projalign_iter_forced(meas)
        d = float(np.abs(own - meas).max())
        p = self._shift_diff()
        self.r.pa_meas.append(d)
        self.r.pa_params.append(p)
        self.log(f"  PA sweep: per-target diff {d:.2e} px, params {p:.1e}, "
                 f"maxerr {eg:.3f}")
        assert d <= SHIFT_TOL, d
        return eg

    def projalign_once(self):
        last = self.projalign_iter()
        for _ in range(1, 10):
            err = self.projalign_iter()
            if err <= last:
                if err < 2.0:
                    break
                last = err

    def correct(self, out_bin=1):
        self.sync()
        al = self.G.correct(out_bin)
        oal = self.O.correct_final(out_bin)
        self.r.aligned_rel = rel_l2(al, oal)
        assert self.r.aligned_rel <= 1e-6, self.r.aligned_rel
        return al

    # ------------------------------------------------------ whole DoIt
    def run(self, align_z, nproj=10, out_bin=1):
        """Process::DoIt's stage sequence (pitch and offset searched)."""
        self.preprocess()
        self.coarse_setup()
        for i in (1, 2, 3):
            self.coarse_iter()
            self.tilt_axis(180.0 / i)
        self.find_offset()
        self.apply_offset()
        self.projalign_setup(align_z, nproj)
        self.projalign_once()
        for i in (1, 2, 3, 4):
            self.tilt_axis(20.0 / i)
            if i == 1:
                self.projalign_once()
        self.projalign_once()
        self.correct(out_bin)
        return self.r

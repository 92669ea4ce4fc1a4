"""Orthogonal projections of an instrument and its neutron trace lines.

The drop-in for the reference's Visualisation (mcramp/visualisation.py:10-251,
opened by Instrument.visualise, mcramp.py:317-325): a 2 x 2 figure with the
z-x, z-y and x-y projections and a 3-D view (lab x, z, y on its x, y, z
axes), every drawn component's geometry and scatter outlines (their
lines(), rotated by frame_rotate and moved to the component position), and
the trace lines of the last traced runs in red.  With controls, span
selectors set the x, y and z ranges and a list of component names centres
the view on one.  Host-side matplotlib only; nothing here runs on the GPU.
"""
import numpy as np

# the three projections: (name, lab axis drawn horizontally, lab axis drawn vertically)
PROJECTIONS = (("zx", 2, 0), ("zy", 2, 1), ("xy", 0, 1))
# lab axis -> axis of the 3-D view (lab y is drawn upwards)
_AXIS3D = {0: "x", 2: "y", 1: "z"}


def component_outlines(comp):
    """Lab-frame outlines of one component: its geometry's then its scatter
    part's lines(), rotated by the component's Euler angles
    (frame_rotate, mcramp.py:457-476) and offset by its position."""
    from .instrument import frame_rotate
    out = []
    for part in (comp.geom, comp.scat):
        pts = frame_rotate(part.lines(), comp.rot)
        out.append([np.add(pts[k], comp.pos[k]) for k in range(3)])
    return out


class Visualisation:
    """Figure of an instrument (mcramp/visualisation.py:10-251)."""

    def __init__(self, inst, controls=True, xlim=None, ylim=None, zlim=None, focus=None,
                 style="dark_background", **kwargs):
        import matplotlib.pyplot as plt
        from mpl_toolkits.mplot3d import Axes3D  # noqa: F401  (registers the 3-D projection)
        plt.style.use(style)
        self._plt = plt
        self.inst = inst
        self.controls = controls
        self.offset = np.zeros(3)
        self.ranges = np.ones(3)
        self.fig = plt.figure(**kwargs)
        self.axes = {name: self.fig.add_subplot(2, 2, k + 1) for k, (name, _, _) in enumerate(PROJECTIONS)}
        self.ax3d = self.fig.add_subplot(2, 2, 4, projection="3d")
        self._spans = {}
        if controls:
            self._add_controls()
        self._draw()
        if focus:
            self.focus(focus)
        for axis, lim in ((0, xlim), (1, ylim), (2, zlim)):
            if lim:
                self.set_limits(axis, lim[0], lim[1])

    # ---- drawing ---------------------------------------------------------
    def _labels(self):
        for name, h, v in PROJECTIONS:
            self.axes[name].set_xlabel("xyz"[h] + " [m]")
            self.axes[name].set_ylabel("xyz"[v] + " [m]")
        self.ax3d.set_xlabel("x [m]")
        self.ax3d.set_ylabel("z [m]")
        self.ax3d.set_zlabel("y [m]")

    def _plot(self, xyz, *fmt, **kw):
        o = self.offset
        p = [np.add(xyz[k], o[k]) for k in range(3)]
        for name, h, v in PROJECTIONS:
            self.axes[name].plot(p[h], p[v], *fmt, **kw)
        self.ax3d.plot(p[0], p[2], p[1], *fmt, **kw)

    def _draw(self):
        for ax in list(self.axes.values()) + [self.ax3d]:
            ax.clear()
        for comp in self.inst.kernel_refs:
            if comp.vis:
                for xyz in component_outlines(comp):
                    self._plot(xyz)
        for block in self.inst.blocks:
            for line in block.trace_lines:
                self._plot(line, "r-", linewidth=1)
        self._labels()
        self._update_ranges()

    def _update_ranges(self):
        """Current extent of each lab axis; the span selectors cover it with
        half of it spare on both sides."""
        zx, xy = self.axes["zx"], self.axes["xy"]
        lims = (zx.get_ylim(), xy.get_ylim(), zx.get_xlim())
        self.ranges = np.array([hi - lo for lo, hi in lims])
        for axis, ax in self._spans.items():
            lo, hi = lims[axis]
            r = self.ranges[axis]
            ax.set_xlim(lo - 0.5 * r, hi + 0.5 * r)

    # ---- view ------------------------------------------------------------
    def set_limits(self, axis, vmin, vmax):
        """Show lab axis `axis` (0 x, 1 y, 2 z) over [vmin, vmax] in every
        view that draws it (an empty range is ignored, as a click is)."""
        if vmin == vmax:
            return
        for name, h, v in PROJECTIONS:
            if h == axis:
                self.axes[name].set_xlim(vmin, vmax, emit=False, auto=False)
            if v == axis:
                self.axes[name].set_ylim(vmin, vmax, emit=False, auto=False)
        getattr(self.ax3d, "set_%slim" % _AXIS3D[axis])(vmin, vmax, emit=False, auto=False)

    def focus(self, comp_name):
        """Centre the view on a component: draw everything shifted by minus
        its position, then show a tenth of the x and z extent around it."""
        for comp in self.inst.kernel_refs:
            if comp.name == comp_name:
                self.offset = -np.asarray(comp.pos, dtype=np.float64)
                self._draw()
                self.set_limits(0, -0.1 * self.ranges[0], 0.1 * self.ranges[0])
                self.set_limits(2, -0.1 * self.ranges[2], 0.1 * self.ranges[2])
                self.fig.canvas.draw_idle()

    def _add_controls(self):
        import matplotlib.widgets as wid
        plt = self._plt
        plt.subplots_adjust(left=0.05, right=0.7)
        props = dict(facecolor="blue", alpha=0.5)
        self._widgets = []
        for axis, (y_span, y_label) in enumerate(((0.3, 0.25), (0.2, 0.15), (0.1, 0.05))):
            ax = plt.axes([0.72, y_span, 0.25, 0.03])
            ax.set_yticks([])
            self._spans[axis] = ax
            self._widgets.append(wid.SpanSelector(ax, lambda lo, hi, a=axis: self.set_limits(a, lo, hi),
                                                  "horizontal", props=props))
            label = plt.axes([0.72, y_label, 0.25, 0.03])
            label.text(0, 0, "xyz"[axis] + "lim")
            label.axis("off")
        names = tuple(c.name for c in self.inst.kernel_refs)
        if names:
            ax = plt.axes([0.72, 0.35, 0.25, 0.63], facecolor="grey")
            radio = wid.RadioButtons(ax, names)
            radio.on_clicked(self.focus)
            self._widgets.append(radio)

    def show(self):
        self._plt.show()

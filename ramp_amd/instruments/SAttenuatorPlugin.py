"""A user-written scatter plugin, as a user would add one: a module named
like the component, holding a ScatterPlugin subclass with HIP-C source.
It moves the neutron to the exit of its geometry, multiplies the weight by
a random factor in [f, 1] and kills the neutron with probability q - only
IEEE-exact arithmetic and the neutron's own RNG stream, so that its host
build reproduces the GPU bit for bit."""
from ramp_amd.plugins import ScatterPlugin

SOURCE = r"""
RAMP_SCATTER void attenuate(RampNeutron *n, RampIntersection *is, uint32_t *iidx, uint32_t comp_idx,
                            const float *params, uint32_t tid)
{
    if (*iidx != comp_idx) return;
    if (n->die > 0.0f) return;
    float u = ramp_rand(n, tid);
    n->p *= params[0] + (1.0f - params[0]) * u;
    if (ramp_rand(n, tid) < params[1]) n->die = 1.0f;
    n->x = is->x2; n->y = is->y2; n->z = is->z2;
    n->tof += is->t2;
    *iidx = 0;
    n->sc = (float)comp_idx;
    ramp_isect_reset(is);
}
"""


class SAttenuatorPlugin(ScatterPlugin):
    entry = "attenuate"
    source = SOURCE

    def __init__(self, f=1.0, q=0.0, idx=0, ctx=None, **kwargs):
        super().__init__(params=[f, q], idx=idx)

"""Edge cases of the run on the GPU: an instrument holding only its source
runs the source segment alone (the reference runs gen_prg and no component
kernel) and leaves the generated neutrons, equal to the oracle's."""
import json

import numpy as np
import pytest

from ramp_amd import Instrument
from tests.parity_util import compare_states

SOURCE_ONLY = {
    "beam": {
        "linear": True,
        "source": {"source": True,
                   "moderator_kernel": {"name": "MFlat", "mod_dim": [0.04, 0.09], "target_dim": [0.03, 0.03],
                                        "target_dist": 2.0, "E_min": 2.0, "E_max": 20.0}},
    }
}


@pytest.mark.gpu
def test_source_only_instrument(tmp_path):
    from oracle import pyoracle
    p = tmp_path / "src.json"
    p.write_text(json.dumps(SOURCE_ONLY))
    print("source-only instrument", flush=True)
    g = Instrument(str(p))
    g.execute(5000, seeds=[3], keep_state=True)
    gs = g.state(5000)
    assert g.stats()["n_segments"] == 1
    g.close()
    o = pyoracle.run(Instrument(str(p)), 5000, [3], keep_state=True)
    d = compare_states(gs, o["state"])
    assert d["dead"] == 0 and d["ctr"] == 0 and d["beyond_rows"] == 0
    assert (gs[0][:, 15] == 0).all() and np.isfinite(gs[0][:, :11]).all()

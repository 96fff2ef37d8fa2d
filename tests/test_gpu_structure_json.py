"""The structure and rank report of a BLR matrix built on the GPU against the
oracle's report of its own construction of the same matrix (write_JSON,
src/util/print.cpp:54-128)."""
import json

import numpy as np
import pytest

from oracle import frank_oracle as fo

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_report_matches_oracle(torch_gpu, fa, tmp_path, precision):
    N, nb, eps = 1024, 4, 1e-8
    print(f"structure json N={N} nb={nb} eps={eps} {precision}", flush=True)
    x = fo.get_sorted_random_vector(N)
    A = fa.BlrMatrix.from_kernel([x], nb, eps=eps, kcap=64)
    if precision == "f32":
        A = A.convert("f32")
    doc = A.to_json()
    ker = lambda r0, nr, c0, nc: fo.laplacend([x], r0, nr, c0, nc)
    ref = fo.to_json(fo.build_blr(ker, N, N, nb, nb, eps=eps, points=[x]))
    for k in ("type", "dim", "abs_pos", "level"):
        assert doc[k] == ref[k]
    # FP32 storage rounds every entry: singular values move by about 1e-7 of the largest
    tol = 10 * eps if precision == "f64" else 1e-6
    for row, rrow in zip(doc["children"], ref["children"], strict=True):
        for a, b in zip(row, rrow, strict=True):
            assert a.keys() == b.keys()
            for k in ("type", "dim", "abs_pos", "level"):
                assert a[k] == b[k]
            if a["type"] == "LowRank":
                assert a["rank"] == b["rank"]
            s, sr = np.array(a["svalues"]), np.array(b["svalues"])
            assert s.shape == sr.shape and np.all(np.diff(s) <= 0)
            np.testing.assert_allclose(s, sr, rtol=0, atol=tol * sr[0])
    path = tmp_path / "blr.json"
    A.write_json(str(path))
    text = path.read_text()
    assert "\n" not in text and text.startswith('{"abs_pos":[0,0],"children":')
    assert json.loads(text) == json.loads(json.dumps(doc))

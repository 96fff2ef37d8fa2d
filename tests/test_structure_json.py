"""Structure and rank report (write_JSON, src/util/print.cpp:54-128) on the CPU.

The oracle's to_json and the library's host formatter (frank_amd.blr.
structure_json, fed through tile accessors) are both checked against a
document written out by hand for a 2 x 2 grid whose blocks have known
singular values: a diagonal Dense block, a rank-1 LowRank block with a single
nonzero entry, a Dense block with one nonzero entry, and a rank-2 LowRank block
with permuted identity factors.
"""
import json

import numpy as np
import pytest

from oracle import frank_oracle as fo


def _grid():
    d00 = np.diag([3.0, 2.0, 1.0])
    lr01 = fo.LowRank(np.array([[1.0], [0.0], [0.0]]), np.array([[5.0]]), np.array([[0.0, 1.0]]))
    d10 = np.array([[0.0, 0.0, 4.0], [0.0, 0.0, 0.0]])
    lr11 = fo.LowRank(np.eye(2), np.diag([7.0, 6.0]), np.array([[0.0, 1.0], [1.0, 0.0]]))
    return fo.Hier(2, 2, [d00, lr01, d10, lr11], [(0, 3), (3, 2)], [(0, 3), (3, 2)])


EXPECTED = {
    "type": "Hierarchical", "dim": [5, 5], "abs_pos": [0, 0], "level": 0,
    "children": [
        [{"type": "Dense", "dim": [3, 3], "abs_pos": [0, 0], "level": 1, "svalues": [3.0, 2.0, 1.0]},
         {"type": "LowRank", "dim": [3, 2], "abs_pos": [0, 1], "level": 1, "svalues": [5.0, 0.0], "rank": 1}],
        [{"type": "Dense", "dim": [2, 3], "abs_pos": [1, 0], "level": 1, "svalues": [4.0, 0.0]},
         {"type": "LowRank", "dim": [2, 2], "abs_pos": [1, 1], "level": 1, "svalues": [7.0, 6.0], "rank": 2}],
    ],
}


def _same(doc, ref):
    assert doc.keys() == ref.keys()
    for k in ref:
        if k == "children":
            assert len(doc[k]) == len(ref[k])
            for row, rrow in zip(doc[k], ref[k]):
                assert len(row) == len(rrow)
                for a, b in zip(row, rrow):
                    _same(a, b)
        elif k == "svalues":
            np.testing.assert_allclose(doc[k], ref[k], rtol=0, atol=1e-14)
        else:
            assert doc[k] == ref[k], k


def _accessors(H):
    from frank_amd import _lib as L

    def tile_info(i, j):
        b = H[i, j]
        m, n = H.row_ranges[i][1], H.col_ranges[j][1]
        if b is None:
            return L.TILE_EMPTY, 0, m, n
        if isinstance(b, fo.LowRank):
            return L.TILE_LOWRANK, b.rank, m, n
        return L.TILE_DENSE, 0, m, n

    def tile(i, j):
        b = H[i, j]
        return (b.U, b.S, b.V) if isinstance(b, fo.LowRank) else b

    return (H.nb_r, H.nb_c), tile_info, tile


def test_oracle_to_json_matches_hand_written_document():
    _same(fo.to_json(_grid()), EXPECTED)


def test_host_formatter_matches_hand_written_document_and_oracle():
    from frank_amd.blr import structure_json
    H = _grid()
    doc = structure_json(*_accessors(H))
    _same(doc, EXPECTED)
    _same(doc, fo.to_json(H))
    # compact, sorted keys, as write_JSON's stream output of an nlohmann::json (std::map keys)
    text = json.dumps(doc, sort_keys=True, separators=(",", ":"))
    assert text.startswith('{"abs_pos":[0,0],"children":[[{"abs_pos":[0,0],"dim":[3,3],"level":1,')
    _same(json.loads(text), EXPECTED)


def test_empty_block_has_no_report():
    from frank_amd.blr import structure_json
    H = _grid()
    H[1, 0] = None
    with pytest.raises(TypeError):
        fo.to_json(H)
    with pytest.raises(TypeError):
        structure_json(*_accessors(H))


def test_zero_rank_block_reports_zero_singular_values():
    from frank_amd.blr import structure_json
    H = _grid()
    H[0, 1] = fo.LowRank(np.zeros((3, 0)), np.zeros((0, 0)), np.zeros((0, 2)))
    doc = structure_json(*_accessors(H))
    node = doc["children"][0][1]
    assert node["rank"] == 0 and node["svalues"] == [0.0, 0.0]
    _same(doc, fo.to_json(H))

"""Reference KDTreeData layout (the oracle's restatement of kernel_TraversalDownPath):
a hand-derived known-answer stream from the documented format (KDKernelDefs.h:76-117)
and a decode round trip back to the packed tree on oracle-built trees."""

import numpy as np

from mnrt_amd import scene as S


def _fbits(x):
    return int(np.float32(x).view(np.uint32))


def test_export_known_answer(orc):
    # root: split y at 0.25 -> left leaf {2, 5}; right: split x at -0.5 -> leaf {1}, empty leaf
    L = 3 << 30
    nodes = np.array([_fbits(0.25), (1 << 30) | 2, 2, L | 0, _fbits(-0.5), (0 << 30) | 4, 1, L | 2, 0, L | 3],
                     np.uint32)
    refs = np.array([2, 5, 1], np.uint32)
    root = np.array([-1, -1, -1, 1, 1, 1], np.float32)
    X = orc.export_preorder(nodes, refs, root, num_leaves=3)
    F = 0x80000000
    want = [0, (1 << 30) | 7, _fbits(0.25),  # node 0 at word 0, right child (node 2) at word 7
            1 | F, 2, 2, 5,                    # node 1: leaf, 2 elements
            2, (0 << 30) | 13, _fbits(-0.5),   # node 2 at word 7, right child (node 4) at word 13
            3 | F, 1, 1,                       # node 3: leaf, 1 element
            4 | F, 0]                          # node 4: empty leaf
    assert X["tree"].tolist() == want
    assert X["addr"].tolist() == [0, 3, 7, 10, 13]
    assert X["num_elems"].tolist() == [3, 2, 1, 1, 0]
    assert X["left"].tolist() == [1, 0, 3, 0, 0] and X["right"].tolist() == [2, 0, 4, 0, 0]
    boxes = {0: ([-1, -1, -1], [1, 1, 1]), 1: ([-1, -1, -1], [1, 0.25, 1]), 2: ([-1, 0.25, -1], [1, 1, 1]),
             3: ([-1, 0.25, -1], [-0.5, 1, 1]), 4: ([-0.5, 0.25, -1], [1, 1, 1])}
    for i, (lo, hi) in boxes.items():
        lo, hi = np.array(lo, np.float64), np.array(hi, np.float64)
        np.testing.assert_allclose(X["extent"][i, :3], (lo + hi) / 2, rtol=0, atol=1e-7)
        np.testing.assert_allclose(X["extent"][i, 3], np.linalg.norm(hi - lo) / 2, rtol=1e-7)


def _decode(tree, n):
    """Parses the reference stream back into packed node words and leaf references."""
    nodes = np.zeros(2 * n, np.uint32)
    refs = []
    order = []

    def walk(a):
        w = int(tree[a])
        i = w & 0x7FFFFFFF
        order.append(i)
        if w & 0x80000000:
            cnt = int(tree[a + 1])
            nodes[2 * i] = cnt
            nodes[2 * i + 1] = (3 << 30) | len(refs)
            refs.extend(int(x) for x in tree[a + 2:a + 2 + cnt])
            return a + 2 + cnt
        info, split = int(tree[a + 1]), int(tree[a + 2])
        ar = info & 0x0FFFFFFF
        end_left = walk(a + 3)
        assert end_left == ar, "right child must follow the left subtree"
        r = int(tree[ar]) & 0x7FFFFFFF
        nodes[2 * i] = split
        nodes[2 * i + 1] = ((info >> 30) << 30) | r
        return walk(ar)

    end = walk(0)
    assert end == len(tree)
    assert order == list(range(n)), "stream order is preorder"
    return nodes, np.array(refs, np.uint32)


def test_export_round_trip(orc):
    for sc in (S.scene_small(subdiv=2), S.scene_sized(3000, seed=5)):
        t = orc.build(0, sc.v0, sc.v1, sc.v2)
        X = orc.export_preorder(t.nodes, t.refs, t.root, t.num_leaves)
        nodes, refs = _decode(X["tree"], t.num_nodes)
        assert np.array_equal(nodes, t.nodes)
        assert np.array_equal(refs, t.refs)
        leaf = (t.nodes[1::2] >> 30) == 3
        assert X["num_elems"][0] == t.refs.size
        assert np.array_equal(X["num_elems"][leaf], t.nodes[0::2][leaf])
        # every inner node's count is the sum of its children's
        inner = np.nonzero(~leaf)[0]
        assert np.array_equal(X["num_elems"][inner], X["num_elems"][inner + 1] + X["num_elems"][X["right"][inner]])

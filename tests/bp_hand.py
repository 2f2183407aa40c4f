"""Known answer for BP on three vertices, derived by hand."""
import numpy as np

# A 3-vertex known answer worked out by hand from numa-BP.C:112-125 and the loop at
# numa-BP.C:280-316: edges 0->2 (e=0), 1->2 (e=1), 2->0 (e=2); W_e[k][l] is stored at
# 4e+2k+l and is asymmetric, so a transposed W, the source's potential or product in
# place of the destination's, a missing reset of active vertices, or a reset of an
# inactive one each change the numbers.  Vertex 2 has two in-edges, vertex 1 none.
#   iteration 1 (all active): message into 2 over e: b[k] = sum_l phi2[l] W_e[k][l] P2[l]
#     with phi2*P2 = (2*1, 3*2) = (2, 6): e=0 -> (1*2+2*6, 3*2+4*6) = (14, 30),
#     e=1 -> (2*2+1*6, 1*2+3*6) = (10, 20); P2 = (14*10, 30*20) = (140, 600).
#     into 0 over e=2: phi0*P0 = (1*1, 2*3) = (1, 6) -> (1+6, 2+6) = (7, 8).
#     vertex 1 has no in-edge: the reset value (1, 1).
#   iteration 2 (frontier {0, 2}, the vertices reached in iteration 1): vertex 1 is
#     inactive and keeps the other buffer's content, the initial product (2, 2);
#     into 2 over e=0 only: phi2*P2 = (280, 1800) -> (280+3600, 840+7200) = (3880, 8040);
#     into 0 over e=2: phi0*P0 = (7, 16) -> (7+16, 14+16) = (23, 30).
BP_HAND_OFF = np.array([0, 1, 2, 3], np.int64)
BP_HAND_ADJ = np.array([2, 2, 0], np.int32)
BP_HAND_VPOT = np.array([[1, 2], [3, 5], [2, 3]], np.float32)
BP_HAND_EPOT = np.array([1, 2, 3, 4, 2, 1, 1, 3, 1, 1, 2, 1], np.float32)
BP_HAND_PROD0 = np.array([[1, 3], [2, 2], [1, 2]], np.float64)
BP_HAND_ANSWER = {
    1: np.array([[7, 8], [1, 1], [140, 600]], np.float64),
    2: np.array([[23, 30], [2, 2], [3880, 8040]], np.float64),
}

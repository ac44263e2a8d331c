"""Algorithmic byte model of the tracer kernels (for roofline fractions).

Per unit of work each kernel must move at least these bytes between HBM and
the CUs (SoA path state, BVH2 layout of csrc/bvh.hpp):

  node visit      64 B   one bvh2 node (two child boxes + child refs); 128 B for a
                         bvh4 node (four child boxes + refs, one cache line); 80 B
                         for a bvh8 node (eight 8-bit quantised child boxes)
  primitive test  48 B   one leaf slot (three float4: p0|id, p1, p2)
  extend ray      36 B   queue index 4 + ray origin/direction 24 (read)
                         + hit prim/t 8 (write)
  shade path     264 B   queue 4, hit 8, ray 24, beta 12, value 12, rng 8,
                         pixel/sample/bounce 12 (read); 3 vertex positions 36,
                         3 normals 36, 3 indices 12, entity 48, material 48 (gather);
                         ray 24, beta 12, value 12, rng 8, bounce 4 (write)
  connect record 112 B   flags 4, shadow ray 28 + light id 4, contribution 12, vertex
                         point/normal 24, BSDF direction/value/pdf 28, beta 12: written
                         by shade and read by connect, so counted in both kernels
  connect extra   28 B   connect's queue index 4, value 12 (read) and 12 (write)
  light cast      96 B   two light triangles' vertex positions and indices
"""
S_NODE = 64
S_NODE4 = 128
S_NODE8 = 80
S_PRIM = 48
B_EXTEND_RAY = 36
B_SHADE_PATH = 264
B_CONNECT_REC = 112
B_CONNECT_EXTRA = 28
B_LIGHT_CAST = 96

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def kernel_bytes(stats, width=2):
    """Algorithmic bytes of a render per kernel kind, from an instrumented run's counters
    (width: the BVH the instrumented run walked)."""
    node = {8: S_NODE8, 4: S_NODE4}.get(width, S_NODE)
    ext_nodes = stats["extend_node_visits"]
    ext_prims = stats["extend_prim_tests"]
    con_nodes = stats["node_visits"] - ext_nodes
    con_prims = stats["prim_tests"] - ext_prims
    ext_rays = stats["extension_rays"]
    records = max(stats["shadow_rays"], stats["mis_rays"])
    return {
        "extend": ext_rays * B_EXTEND_RAY + ext_nodes * node + ext_prims * S_PRIM,
        # shade writes the connect record, connect reads it back
        "shade": ext_rays * B_SHADE_PATH + records * B_CONNECT_REC,
        "connect": records * (B_CONNECT_REC + B_CONNECT_EXTRA) + con_nodes * node + con_prims * S_PRIM
        + stats["light_casts"] * B_LIGHT_CAST,
    }


# The guiding network (9 inputs -> 216 sin/cos features -> 5 x 64 ReLU -> 40 head
# outputs): multiply-adds per record of the live (unpadded) matrices.
MLP_MACS_FORWARD = 216 * 64 + 4 * 64 * 64 + 64 * 40          # 32,768
MLP_MACS_DATA_GRAD = 4 * 64 * 64 + 64 * 40                   # layers 1-5 (the encoding needs none)
MLP_FLOP_FORWARD = 2 * MLP_MACS_FORWARD                      # 65,536 per record
# one training record: forward with stored activations, data gradients of layers 1-5,
# weight gradients of all six layers (as many multiply-adds as the forward pass)
MLP_FLOP_TRAIN = 2 * (2 * MLP_MACS_FORWARD + MLP_MACS_DATA_GRAD)  # 168,960 per record

MFMA_FP16_PEAK_TFLOPS = 2500.0  # MI355X dense FP16 matrix peak, 2.5 PFLOP/s

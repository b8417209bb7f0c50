"""dtype contract (cnnclustering _primitive_types.pxd:4-7).

Labels are int64 (np.intp on 64-bit hosts), coordinates float64 on the host
API, consider flags uint8.  Device kernels index points with int32 and store
neighbour-list offsets as int64.
"""

import numpy as np

P_AINDEX = np.intp
P_AVALUE = np.float64
P_AVALUE32 = np.float32
P_ABOOL = np.uint8

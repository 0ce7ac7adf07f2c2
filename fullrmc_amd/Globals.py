"""Type contract of the path (fullrmc Globals.py:29-46): float32 values,
int32 indexes, pi as float32."""
import numpy as np

INT_TYPE = np.int32
FLOAT_TYPE = np.float32
PRECISION = FLOAT_TYPE(1e-5)
PI = FLOAT_TYPE(np.pi)

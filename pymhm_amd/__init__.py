"""MI355X-native MHM local-problem stage (hand-written HIP for gfx950 behind a C ABI).

Modules:

* ``_lib``      ctypes binding of ``libmhm_hip.so`` (``include/mhm_hip.h``)
* ``elements``  reference-element quadrature and Lagrange tables
* ``meshing``   macro/local meshes, nodal numbering, trace coupling
* ``layout``    band ordering, CSR pattern and assembly maps of one local shape
* ``stage``     ``LocalBatch``: resident batched assembly/factor/solve/condense
* ``problems``  providers of the configured workloads
* ``hybrid``    ordered global scatter of condensed blocks
* ``batched``   drop-ins for ``condense_batched`` / ``BatchedFactorization``
* ``serving``   continuous-batching request loop with resident factors
"""

__version__ = "1.0.0"

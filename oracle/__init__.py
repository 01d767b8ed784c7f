"""ORACLE -- CPU FP64 restatement of jaxhps's uniform HPS build/solve.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import, call or link anything in
this package, and only as the checker (or the timed CPU baseline) -- never as
the product path.  The product (``jaxhps_amd``) must fail loudly when its HIP
library is missing; it never falls back to this code.

Modules
-------
spectral       leaf operators and point sets, built independently from
               coordinates and textbook formulas
hps2d          stage-by-stage restatement of the reference's 2D DtN / ItI
               leaf solve, merges, down pass and ItI up pass
hps3d          3D DtN leaf/merge/down pass restatement
global_system  the global HPS discrete system solved densely (cross-check)
interp         HPS grid <-> rectangular grid by explicit Lagrange products
scattering     ItI -> DtN, plane waves, exterior boundary-integral coupling
adaptive2d     adaptive 2D DtN: refinement, compressed merges, down pass
adaptive3d     adaptive 3D DtN: face-panel compression, 8:1 merges, down pass

Pinning (what checks each function beyond re-typing it):
* spectral: textbook D-matrices n=2,3,4 and Chebyshev/Gauss points from the
  reference's own tests; exactness on polynomials; agreement with the
  product's independently written operators.
* hps2d leaf/merge/down: the reference's closed-form accuracy cases at
  atol 1e-12 (polynomials) / 1e-8 (Helmholtz); global_system at 1e-11.
* hps2d up pass / no-source: equality with the with-source path, as the
  reference's ``test_nosource_accuracy.py`` checks.
* hps3d: leaf operators exact on polynomials (derivatives, P, Q); a
  manufactured cubic solution reproduced to rounding through the oct merges
  and down pass; the 3D global system (coordinates only) on tiny inputs.
* interp: exactness on polynomials of degree below the node counts (2D and
  3D, both directions); the first-leaf rule on shared edges by a leaf-wise
  constant field.
* scattering: the Cayley map of a known DtN inverted; normal derivatives
  against finite differences; a q = 0 Helmholtz DtN reproducing plane-wave
  normal derivatives; the D = I/2 closed form; no scattered field at q = 0.
* adaptive2d: identity with hps2d on uniform trees; exact polynomial
  solutions on non-uniform trees; the reference's panel-norm closed forms;
  level restriction of refined trees.
* adaptive3d: identity with hps3d on uniform octrees; a cubic solution exact
  on a non-uniform octree; exact face projections.
"""

from . import spectral, hps2d, hps3d, global_system  # noqa: F401

__all__ = ["spectral", "hps2d", "hps3d", "global_system"]

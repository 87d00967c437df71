"""CPU oracle for the narrow-band SVI erosion path.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy restatement of the
reference algorithm (jwcalder/SVIErosion, files cited per function) used as
the checker of the HIP implementation in ``svierosion_amd``.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; the product path never does.

Every function runs in the dtype of its inputs: float32 inputs follow the
reference's single-precision CPU path, float64 inputs give the FP64 oracle.

Pinning (what checks each part independently of the restatement itself):

* grid sizing              -- closed-form KAT (cube res 64 -> 90^3, h = 1/64)
* cube_plane_fraction      -- FP64 inclusion-exclusion on 20k random planes,
                              degenerate normals, cube volume = 1 exactly
* godunov_inward (WENO5)   -- |grad phi| = 1 for plane / sphere distance fields,
                              polynomial exactness of each candidate stencil
* closest_points, tricubic -- exact projection for planes, cubic-polynomial
                              exactness of the interpolant
* ball_cell_weights        -- sum = 4/3 pi R^3, brute-force subsampling of cut cells
* ball convolution (FFT)   -- brute-force direct sum on small grids; sphere /
                              plane / cube-corner closed forms
* laws                     -- closed forms, n_cap root property
* point_triangle_distance2 -- brute force over densely sampled triangle points
* mesh_to_sdf              -- analytic box SDF, brute force over all triangles
* inside_mask              -- analytic inside tests (box, sphere, torus)
* reinitialize             -- distorted-sphere interface preservation
* simulation               -- sphere radius ODE, exact stop at target volume
"""

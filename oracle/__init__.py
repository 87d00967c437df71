"""CPU oracle for the narrow-band SVI erosion path.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy restatement of the
reference algorithm (jwcalder/SVIErosion, files cited per function) used as
the checker of the HIP implementation in ``svierosion_amd``.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; the product path never does.

Every function runs in the dtype of its inputs: float32 inputs follow the
reference's single-precision CPU path, float64 inputs give the FP64 oracle.

Pinning: what checks each part independently of the restatement itself, and
the test that does it.

tests/test_oracle.py
* grid sizing              -- closed-form KAT (cube res 64 -> 90^3, h = 1/64),
                              next_fast_size against prime factorisation
* cube_plane_fraction      -- FP64 inclusion-exclusion on 20k random planes,
                              degenerate normals, Monte-Carlo counting
* LevelSet.volume          -- ball volume converges to 4/3 pi R^3
* godunov_inward (WENO5)   -- |grad phi| = 1 for plane / sphere distance fields,
                              exactness on quadratic data
* closest_points, tricubic,
  trilinear                -- exact projection onto planes, cubic / linear
                              exactness of the interpolants
* principal_curvatures     -- sphere: k1 = k2 = 1/r
* point_triangle_distance2 -- plane / segment distances (independent formula)
* ball_cell_weights        -- sum = 4/3 pi R^3, Monte-Carlo cut cells
* ball convolution (FFT)   -- brute-force direct sum; sphere closed form at r/h = 8
* SVILaw                   -- n_cap root property, dspeed vs finite differences
* mesh_to_sdf              -- analytic box SDF (< 2e-5 h)
* inside_mask              -- analytic sphere inside test, signed and parity
* marching_tetrahedra      -- level offsets and a displaced vertex measured by
                              the Hausdorff checker

tests/test_oracle_flow.py
* CurvatureLaw             -- closed forms (alpha = 1, 2), diffusivity = sum of
                              the speed's partial derivatives by central
                              differences (alpha < 1 included), regularisation
                              asymptotics, parameter rejections
* LevelSet.reinitialize    -- distorted-sphere interface preservation within
                              the reference's 0.01 h / 0.005 h bounds
* Simulation.run           -- sphere radius ODEs (three SVI laws at the
                              reference's sizes and bounds; four curvature laws
                              incl. the alpha = 1/2 regularised mean law), exact
                              stop at the target volume on the cube (five
                              places, mesh volume ratio within 2e-3)
* inside_mask              -- analytic box and torus inside tests

tests/test_gpu_surface.py pins oracle/surface.py by an analytic ellipsoid
(watertight, oriented, volume within 2e-3).
"""

"""FP64 CPU oracle for MARTINI's source-insertion path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``martini_amd`` imports, links or runs
anything under this folder; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Contents
--------
``martini_oracle``  numpy FP64 restatement of the reference algorithm
                    (projection, channel edges, kernel selection, pruning,
                    kernel integrals, spectra, pixel-loop cube).
``cube_oracle.c``   the same cube (FP64, particle-by-particle) in C, for sizes
                    that the numpy pixel loop cannot finish in seconds.
``cpu_baseline``    the reference's own CPU algorithm restated as it runs
                    (per-pixel O(N) mask, N x Nc spectra, ThreadPool(ncpu)),
                    timed as the CPU baseline.
"""

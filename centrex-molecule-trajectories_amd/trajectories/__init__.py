"""MI355X (gfx950) trajectory engine with the API of otimgren/centrex-molecule-trajectories.

Drop-in modules: trajectories.beamline, trajectories.beamline_elements, trajectories.distributions,
trajectories.molecule, trajectories.trajectory_simulator, trajectories.post_processing,
trajectories.utils (reading saved runs; with trajectories.hdf it needs h5py).  The hot path
(propagation through the beamline) runs in libctraj.so, built from csrc/ with hipcc for gfx950.
"""

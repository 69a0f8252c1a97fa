"""Running a gRASPA deck on the device-resident driver.

run_deck follows the reference's single-box run (axpy.cu:469-641,
main.cpp:354-410): the deck is read, the box and its Ewald sums are set up,
initialization, equilibration and production cycles run on the GPU
(graspa_batch_* with one box), and at the end the energy is recomputed and
checked against the running sum, the statistics are written in the layout of
the reference's output.txt, and the final configuration is written as a
restart file.  Movie snapshots (LAMMPS data files) are written during
production every MoviesEvery cycles, at the first launch boundary at or after
each multiple (a launch is at most moves_per_launch moves).

Several decks run together by putting their boxes into one graspa_batch
(graspa_amd.serve does this for a stream of jobs); this function runs one.
"""
from __future__ import annotations

import os
import time

from . import deck, engine, report, writers


def run_deck(dirpath, out_dir=None, precision=engine.GRASPA_MIXED, device=0, moves_per_launch=2048,
             framework_name=None, capacity=None, log=None):
    """Runs the deck in `dirpath`; files go under `out_dir` (default: the deck
    directory): Output/System_0/output.txt, Restart/System_0/restartfile,
    Movies/System_0/result_<cycle>.data.  Returns the RunRecord."""
    s = deck.read_deck(dirpath)
    out_dir = out_dir or dirpath
    name = framework_name or os.path.basename(os.path.normpath(dirpath))
    box = engine.Box(s, precision=precision, device=device, capacity=capacity)
    box.init_ewald()

    def say(msg):
        if log is not None:
            log(msg)

    if any(c.create_molecules for c in s.components):
        # CreateMolecule_InOneBox (axpy.cu:300-377, main.cpp:338): the
        # requested molecules are created by CBMC insertions in the
        # reference's draw order (RandomSeed) before the cycles start; the
        # energy summary starts from the box after creation
        mc = engine.MonteCarlo(box)
        try:
            for ci, c in enumerate(s.components, start=1):
                if c.create_molecules > 0:
                    say("Component %d, Need to create %d full molecule" % (ci, c.create_molecules))
                    mc.create_molecules(ci, c.create_molecules)
        finally:
            mc.close()
    n_mol0 = [len(box.get_molecules(ci)[1]) // c.molsize for ci, c in enumerate(s.components, start=1)]
    initial = report.recompute_totals(box, s, n_mol0)
    bt = engine.Batch([box], [s.seed], [0], cycle_cap=max(1, s.n_prod))
    t0 = time.perf_counter()

    try:
        for n, label in ((s.n_init, "INITIALIZATION"), (s.n_equil, "EQUILIBRATION")):
            if n > 0:
                say("%s: %d cycles" % (label, n))
                bt.run(n, production=False, moves_per_launch=moves_per_launch)
        say("PRODUCTION: %d cycles" % s.n_prod)
        bt.set_cycles(0, s.n_prod, production=True)
        first = int(bt.progress()[1][0])
        next_movie = s.movies_every if s.movies_every > 0 else None
        next_print = s.print_every if s.print_every > 0 else None
        while True:
            running = bt.launch(moves_per_launch)
            done = int(bt.progress()[1][0]) - first
            if next_print is not None and done >= next_print:
                say("Production cycle %d of %d" % (done, s.n_prod))
                while next_print <= done:
                    next_print += s.print_every
            if next_movie is not None and done >= next_movie:
                bt.sync()
                writers.write_lammps_movie(os.path.join(out_dir, "Movies", "System_0", "result_%d.data" % done), s,
                                           writers.box_molecules(box))
                while next_movie <= done:
                    next_movie += s.movies_every
            if running == 0 or done >= s.n_prod:
                break
        seconds = time.perf_counter() - t0
        rec = report.record_from_batch(bt, 0, box, s, initial, seconds=seconds, framework_name=name,
                                       capacity=capacity)
        os.makedirs(os.path.join(out_dir, "Output", "System_0"), exist_ok=True)
        report.write_output(rec, os.path.join(out_dir, "Output", "System_0", "output.txt"))
        writers.write_restart(os.path.join(out_dir, "Restart", "System_0", "restartfile"), s,
                              writers.box_molecules(box))
        return rec
    finally:
        bt.close()
        box.close()

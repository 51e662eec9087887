"""Controller (pyclaw/controller.py:275-395): the output-time loop around
Solver.evolve_to_time, keeping in-memory frames (keep_copy) and writing each
frame to outdir in output_format (ascii; None writes nothing)."""
import copy
import os

import numpy as np


class Controller(object):
    def __init__(self):
        self.viewable_attributes = ["xdir", "rundir", "outdir", "overwrite", "xclawcmd",
                                    "xclawout", "xclawerr", "runmake", "savecode", "solver",
                                    "keep_copy", "write_aux_init", "write_aux_always",
                                    "output_format", "output_file_prefix", "output_options",
                                    "num_output_times", "output_style", "verbosity"]
        self.solver = None
        self.solution = None
        self.outdir = "./_output"
        self.overwrite = True
        self.keep_copy = False
        self.write_aux_init = False
        self.write_aux_always = False
        self.output_format = "ascii"
        self.output_file_prefix = None
        self.output_options = {}
        self.num_output_times = 10
        self.output_style = 1
        self.nstepout = 1
        self.out_times = np.linspace(0.0, 1.0, self.num_output_times)
        self.tfinal = 1.0
        self.frames = []
        self.verbosity = 0
        self.compute_p = None
        self.compute_F = None
        self.setplot = None
        self.start_frame = 0

    @property
    def t(self):
        return None if self.solution is None else self.solution.t

    def check_validity(self):
        valid, reason = self.solver.is_valid()
        if not valid:
            raise Exception("Solver is not valid: %s" % reason)

    def run(self):
        if self.solver is None or self.solution is None:
            raise Exception("To run, a Controller must have a Solver and a Solution.")
        self.start_frame = self.solution.start_frame
        if not self.solver._is_set_up:
            self.solver.setup(self.solution)
            self.solver.dt = self.solver.dt_initial
        self.check_validity()
        if self.output_style == 1:
            output_times = np.linspace(self.solution.t, self.tfinal, self.num_output_times + 1)
        elif self.output_style == 2:
            output_times = self.out_times
        elif self.output_style == 3:
            output_times = np.ones((self.num_output_times + 1 - self.start_frame))
        else:
            raise Exception("Invalid output style %s" % self.output_style)
        if len(output_times) == 0:
            print("No valid output times; halting.")
            return None
        frame = self.start_frame
        if self.keep_copy:
            self.frames.append(copy.deepcopy(self.solution))
        if self.output_format is not None:
            # controller.py:329-350: refuse to overwrite unless allowed, then frame 0
            if os.path.exists(self.outdir) and self.overwrite is False:
                raise Exception("Refusing to overwrite existing output data. \n"
                                "Either delete/move the directory or set "
                                "controller.overwrite=True.")
            self._write(frame, self.write_aux_always or self.write_aux_init)
        status = None
        for t in output_times[1:]:
            if self.output_style < 3:
                status = self.solver.evolve_to_time(self.solution, t)
            else:
                for _ in range(self.nstepout):
                    status = self.solver.evolve_to_time(self.solution)
            frame += 1
            if self.keep_copy:
                self.frames.append(copy.deepcopy(self.solution))
            if self.output_format is not None:
                self._write(frame, self.write_aux_always)
        self.solution.__dict__["_start_frame"] = len(self.frames)
        return status

    def _write(self, frame, write_aux):
        self.solution.write(frame, self.outdir, self.output_format, self.output_file_prefix,
                            write_aux, self.output_options)

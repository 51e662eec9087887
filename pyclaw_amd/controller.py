"""Controller (pyclaw/controller.py:275-395): the output-time loop around
Solver.evolve_to_time, keeping in-memory frames.  Writing frames to disk is not
part of this package; output_format is accepted and ignored."""
import copy

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
        if self.keep_copy:
            self.frames.append(copy.deepcopy(self.solution))
        status = None
        for t in output_times[1:]:
            if self.output_style < 3:
                status = self.solver.evolve_to_time(self.solution, t)
            else:
                for _ in range(self.nstepout):
                    status = self.solver.evolve_to_time(self.solution)
            if self.keep_copy:
                self.frames.append(copy.deepcopy(self.solution))
        self.solution.__dict__["_start_frame"] = len(self.frames)
        return status

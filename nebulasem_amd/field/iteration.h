// Time and refinement loops of the apps (src/solvers/iteration.h): the
// step range of one refinement pass from the controls, fields read at its
// start and written every write_interval steps.  Refinement between passes
// is not part of this build (one pass when amr_step is 0).
#ifndef NSEM_FIELD_ITERATION_H
#define NSEM_FIELD_ITERATION_H

#include <iostream>

#include "field.h"

class Iteration {
    Int starti, endi, i, n_deferred, idf;
    bool skip;

   public:
    Iteration(int step) {
        skip = false;
        if (step < 0) {
            skip = true;
            step = 0;
        }
        starti = Controls::write_interval * step + 1;
        endi = Controls::write_interval * (step + Controls::amr_step);
        if (endi > Controls::end_step) endi = Controls::end_step;
        n_deferred = Controls::n_deferred;
        i = starti;
        Controls::current_step = i;
        idf = 0;
        if (MP::printOn) std::cout << "--------------------------------------------\n";
        Mesh::read_fields(step);
        if (MP::printOn) {
            std::cout << "--------------------------------------------\n";
            MP::printH("Starting iterations.\n");
        }
    }
    bool start() { return i == starti; }
    bool end() {
        if (skip) {
            Mesh::write_fields(0);
            return true;
        }
        if (i > endi) return true;
        if (MP::printOn && idf == 0) {
            if (Controls::state == Controls::STEADY) MP::printH("Step %d\n", i);
            else MP::printH("Time %f\n", i * Controls::dt);
        }
        return false;
    }
    void next() {
        idf++;
        if (idf <= n_deferred) return;
        idf = 0;
        MP::printOn = (MP::host_id == 0 && (MP::hasElapsed(Controls::print_time) || i == Controls::end_step - 1));
        if ((i % Controls::write_interval) == 0) Mesh::write_fields(i / Controls::write_interval);
        i++;
        Controls::current_step = i;
    }
    Int get_step() { return i; }
};

class AmrIteration {
    int starti, endi, i;

   public:
    AmrIteration(bool skip_initial = false) {
        starti = Controls::start_step / Controls::write_interval - (skip_initial ? 0 : Controls::amr_step);
        endi = Controls::end_step / Controls::write_interval;
        if (endi < 1) endi = 1;
        if (!Controls::amr_step) Controls::amr_step = endi;
        i = starti;
        Mesh::LoadMesh(i < 0 ? 0 : i);
    }
    bool start() { return i == starti; }
    bool end() { return i >= endi; }
    void next() {
        i += Controls::amr_step;
        if (i < endi) {
            Prepare::refineMesh(i);
            Mesh::LoadMesh(i);
        }
    }
    int get_step() { return i; }
};

#endif

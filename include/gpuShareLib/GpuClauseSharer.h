/***************************************************************************************
 The interface declared here is GpuShareSat's solver-facing API
 (gpuShareLib/GpuClauseSharer.h), GpuShareSat -- Copyright (c) 2020, Nicolas Prevot,
 distributed under the MIT licence reproduced below.  Declarations, their order and
 their inline defaults are kept identical so that solver code compiled against the
 original header links against this library unchanged (same vtable layout).

Permission is hereby granted, free of charge, to any person obtaining a copy of this software and
associated documentation files (the "Software"), to deal in the Software without restriction,
including without limitation the rights to use, copy, modify, merge, publish, distribute,
sublicense, and/or sell copies of the Software, and to permit persons to whom the Software is
furnished to do so, subject to the following conditions:

The above copyright notice and this permission notice shall be included in all copies or
substantial portions of the Software.

THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS OR IMPLIED, INCLUDING BUT
NOT LIMITED TO THE WARRANTIES OF MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND
NONINFRINGEMENT. IN NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM,
DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR OTHERWISE, ARISING FROM, OUT
OF OR IN CONNECTION WITH THE SOFTWARE OR THE USE OR OTHER DEALINGS IN THE SOFTWARE.
 **************************************************************************************************/
#ifndef GpuClauseSharer_h
#define GpuClauseSharer_h
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>

namespace GpuShare {

// Options of the sharer.  Negative values select the defaults.
struct GpuClauseSharerOptions {
    // Guideline for the number of GPU workgroups (-1: from the device).
    int gpuBlockCountGuideline;
    // Guideline for threads per workgroup (-1: default).
    int gpuThreadsPerBlockGuideline;
    // Minimum duration of one gpuRun() call in microseconds (-1: default 50).
    int minGpuLatencyMicros;
    // 0: quiet, 1: some logging.
    int verbosity;
    // Clause activity decay in (0, 1) (negative: default 0.99999).
    double clauseActivityDecay;
    // Time some operations and expose them as statistics.
    bool quickProf;
    // Initial hit capacity hint per category (-1: default).  This library
    // never drops hits; the value only sizes the first buffer.
    int initReportCountPerCategory;
    // Largest single page-locked host allocation in bytes (-1: from the device).
    int maxPageLockedMemory;

    GpuClauseSharerOptions() {
        gpuBlockCountGuideline = -1;
        gpuThreadsPerBlockGuideline = -1;
        minGpuLatencyMicros = -1;
        verbosity = 1;
        clauseActivityDecay = -1;
        quickProf = true;
        initReportCountPerCategory = -1;
        maxPageLockedMemory = -1;
    }
};

enum GlobalStats {
#define X(v) v,
#include "GlobalStats.h"
#undef X
};

enum OneSolverStats {
#define X(v) v,
#include "OneSolverStats.h"
#undef X
};

class GpuClauseSharer {

    public:

    /* Called from one thread (or under the caller's lock) */

    // One check run: ships new clauses and assignments, checks every clause
    // against every sent assignment, and queues the hits for the solvers.
    // Hits of an assignment may only be delivered by a later call.
    virtual void gpuRun() {}

    // Deletes about half of the GPU clauses (the least active ones).
    virtual void reduceDb() {}

    // Clauses ever added, deleted or not.
    virtual long getAddedClauseCount() {return 0;}

    virtual long getAddedClauseCountAtLastReduceDb() {return 0;}

    virtual bool hasRunOutOfGpuMemoryOnce() {return false;}

    virtual void getGpuMemInfo(size_t &free, size_t &total) {}

    virtual int getGlobalStatCount() {return 0;}

    virtual long getGlobalStat(GlobalStats stat) {return 0;}

    virtual void writeClausesInCnf(FILE *file) {}

    virtual void setVarCount(int newCount) {}

    // solverId: the solver the clause comes from (-1 if none); that solver is
    // not sent the clause back for the assignments it had already sent.
    // Returns the clause id, or -1 if the clause is rejected.
    virtual long addClause(int solverId, int *lits, int count) {return -1; }

    /* Not thread safe with any other method */
    virtual void setCpuSolverCount(int count) {}

    /* Thread safe */

    virtual const char* getOneSolverStatName(OneSolverStats stat) {return NULL; }

    virtual const char* getGlobalStatName(GlobalStats stat) {return NULL; }

    virtual int getOneSolverStatCount() { return 0;}

    /* For a given solver id, always called from the same thread (or under a lock) */

    // Sets the literals in the solver's current assignment; all or nothing.
    // Does not take ownership of lits.
    virtual bool trySetSolverValues(int cpuSolverId, int *lits, int count) {return false; }

    // Unsets the literals' variables in the solver's current assignment.
    virtual void unsetSolverValues(int cpuSolverId, int *lits, int count) { }

    // Sends the solver's current assignment to be checked.  Returns its id
    // (increasing), or -1 if no slot is free.
    virtual long trySendAssignment(int cpuSolverId) {return -1; }

    // Pops one clause reported to the solver.  lits stays valid until the next
    // pop for this solver.  A clause is not reported twice to a solver while
    // that solver keeps it; it can be reported again after being deleted.
    virtual bool popReportedClause(int cpuSolverId, int* &lits, int &count, long &gpuClauseId) {return false; }

    // The first assignment id whose hits have not all been delivered.
    virtual long getLastAssigAllReported(int cpuSolverId) { return 0;}

    // The solver's current assignment as last set (0 true, 1 false, 2 undef).
    virtual void getCurrentAssignment(int cpuSolverId, uint8_t* assig) { }

    /* Any thread; may be slightly stale off the solver's own thread */
    virtual long getOneSolverStat(int cpuSolverId, OneSolverStats stat) { return 0; }

    virtual ~GpuClauseSharer() { }

};

GpuClauseSharer* makeGpuClauseSharerPtr(GpuClauseSharerOptions opts, std::function<void (const std::string &str)> logFunc);
}

#endif

// X-macro list of the per-solver statistic names of GpuShare::GpuClauseSharer,
// in the order of gpuShareLib/OneSolverStats.h of GpuShareSat (part of the
// API).  No include guard on purpose.
X(varUpdatesSentToGpu)
X(assigsSentToGpu)
X(failuresToFindAssig)
X(reportedClauses)
X(reportedClausesUnit)
X(reportedClausesBinary)

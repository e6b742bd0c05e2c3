// X-macro list of the global statistic names of GpuShare::GpuClauseSharer,
// in the order of gpuShareLib/GlobalStats.h of GpuShareSat (part of the API:
// the enum values and names must match).  No include guard on purpose.
X(gpuClauses)
X(gpuClauseLengthSum)
X(gpuClausesAdded)
X(gpuRuns)
// one per clause and run, whatever the number of assignments
X(clauseTestsOnGroups)
// one per (clause, solver with assignments) and run
X(clauseTestsOnAssigs)
X(totalAssigClauseTested)
X(gpuReduceDbs)
X(gpuReports)
X(timeSpentTestingClauses)
X(timeSpentFillingAssigs)
X(timeSpentFillingReported)
X(timeSpentReduceGpuDb)

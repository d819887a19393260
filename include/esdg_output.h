/*
 * esdg_output.h -- Tecplot ASCII output of the reference's plot files, and (esdg_solver.h) the
 * binary checkpoint of a solver.
 *
 * Replaces src/Plots.cpp: each writer produces the file the reference's function of the same
 * purpose writes, byte for byte (title, variable list, one POINT zone of ngl x ngl nodes per
 * element, nodes eta outer / xi inner, numbers in the C++ stream's default format).  The
 * reference composes the file names itself (movie/plot_<n>.tec, movie/Arrivals.tec,
 * movie/MaximumElevations.tec, movie/visc_<n>.tec); here the caller passes the path.  Host
 * arrays in the layouts of esdg_setup.h.  0 on success, negative on error
 * (esdg_setup_last_error()).  Part of libesdg_setup.
 */
#ifndef ESDG_OUTPUT_H
#define ESDG_OUTPUT_H

#ifdef __cplusplus
extern "C" {
#endif

/* PlotSolution (src/Plots.cpp:20-107): plot_var 1 writes H = h + b, hu, hv, b; plot_var 2 the free
 * surface h + b - h_0, hu, hv and b - h_0. */
int esdg_plot_solution(const char* path, int Nelem, int ngl, int plot_var, const double* x, const double* y,
                       const double* q, const double* b, double h_0);
/* PlotArrivalTimings (src/Plots.cpp:151-189): the ArrivalTimings buffer per node. */
int esdg_plot_arrival_times(const char* path, int Nelem, int ngl, const double* x, const double* y,
                            const double* arrival_timings);
/* PlotMaximumElevation (src/Plots.cpp:191-231): the MaximumElevation buffer per node. */
int esdg_plot_maximum_elevation(const char* path, int Nelem, int ngl, const double* x, const double* y,
                                const double* maximum_elevation);
/* PlotViscoseParameter (src/Plots.cpp:349-399): the element viscosity at every node of the
 * element (the file ends with the reference's trailing "Test"). */
int esdg_plot_viscosity(const char* path, int Nelem, int ngl, const double* x, const double* y, const double* eps);
/* PlotEntropy (src/Plots.cpp:401-420): total entropy against time, n records. */
int esdg_plot_entropy(const char* path, int n, const double* times, const double* total_entropy);
/* PlotMass (src/Plots.cpp:438-457): relative mass change (m_i - m_0) / m_0 against time. */
int esdg_plot_mass(const char* path, int n, const double* times, const double* total_mass);
/* PlotTimeSeries (src/Plots.cpp:474-517): gauge records, one line per time with the surface
 * elevation at each gauge.  series is [ngauges][n]; names label the columns.  With the reference's
 * seven ports ("Chennai", "Tuticorin", "Visakhapatnam", "Paradip", "Kochi", "Mormugao", "Okha") the
 * file is the reference's byte for byte; the samples come from esdg_gauge_sample. */
int esdg_plot_time_series(const char* path, int n, const double* times, int ngauges, const char* const* names,
                          const double* series);
/* Not provided: PlotFriction / PlotViscosity / PlotSolutionWithExact, which write fields outside
 * this path. */

#ifdef __cplusplus
}
#endif
#endif

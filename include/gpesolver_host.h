/*
 * gpesolver_host.h -- C ABI of the host layer (parameters file, set-up,
 * orchestration) that sits above gpehip.h.  Same conventions: 0 = success,
 * non-zero = failure with a message copied into err (when given).
 */
#ifndef GPESOLVER_HOST_H
#define GPESOLVER_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* FileManager::load_all_v1 + check_params on a gpe_params.json text
 * (src/filemanager/file_manager.cpp:216-235,436-548).  Touches no device. */
int gpes_validate_params_json(const char *json_text, char *err, int errlen);

/* SimulationManager::initialize + run_simulation (src/manager/sim_manager.cpp:16-27,
 * main.cpp:4-16) with gpe_params.json and all outputs in `workdir`.
 * it_tol > 0 stops imaginary time once |dE/N| between two 100-step energy
 * checks falls below it_tol (relative).  Writes the final E/N (hartree) and
 * the number of IT steps run when the pointers are non-NULL. */
int gpes_run(const char *workdir, double it_tol, double *energy_per_atom, long *it_steps, char *err, int errlen);

/* The reference's on-disk formats for the grid of DIR/gpe_params.json:
 * DIR/NAME.gpe.bin and DIR/NAME.gpe.dat (psi, interleaved complex128 in
 * StdMat3D order), DIR/POTE_NAME.gpe.dat (V) and DIR/energy.gpe.dat
 * (n_energies records of e_kin, e_pot, e_int, e_ext, e_bmf; the total is
 * their sum).  FileManager::save_to_binary_file, save_to_text_file,
 * save_pote_to_text_file, save_energies (src/filemanager/file_manager.cpp:
 * 81-112, 251-288, 387-434).  Touches no device. */
int gpes_write_files(const char *dir, const char *name, const double *psi, const char *pote_name, const double *V,
                     const double *energies, int n_energies, char *err, int errlen);

#ifdef __cplusplus
}
#endif
#endif

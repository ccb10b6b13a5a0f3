// rms_main_pgse -- drop-in for the PGSE engine hpc_code/lib/main_PGSE.cu.
int run_cli(int argc, char **argv, bool pgse);
int main(int argc, char **argv) { return run_cli(argc, argv, true); }

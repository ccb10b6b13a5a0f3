// rms_main -- drop-in for the narrow-pulse engine hpc_code/lib/main.cu.
int run_cli(int argc, char **argv, bool pgse);
int main(int argc, char **argv) { return run_cli(argc, argv, false); }

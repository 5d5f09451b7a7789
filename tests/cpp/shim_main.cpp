// Standalone program around the drop-in driver (shim_drive.cpp).
extern "C" int nbx_shim_drive(int argc, char **argv);

int main(int argc, char **argv) { return nbx_shim_drive(argc, argv); }

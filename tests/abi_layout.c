/* Prints sizeof/offsetof of the drop-in structs of include/game_abi.h, one
   "Struct member offset" line each (compiled by tests/test_abi.py). */
#include <stddef.h>
#include <stdio.h>
#include "game_abi.h"
#define P(S, M) printf("%s %s %ld\n", #S, #M, (long)offsetof(S, M))
int main(void)
{
    printf("Grid sizeof %ld\nState sizeof %ld\nDiagnostics sizeof %ld\nForcings sizeof %ld\n"
           "Irreversible_quantities sizeof %ld\nDualgrid sizeof %ld\nConfig sizeof %ld\n",
           (long)sizeof(Grid), (long)sizeof(State), (long)sizeof(Diagnostics), (long)sizeof(Forcings),
           (long)sizeof(Irreversible_quantities), (long)sizeof(Dualgrid), (long)sizeof(Config));
    P(Grid, no_of_oro_layers); P(Grid, normal_distance); P(Grid, volume); P(Grid, area); P(Grid, z_scalar);
    P(Grid, z_vector); P(Grid, gravity_potential); P(Grid, gravity_m); P(Grid, slope); P(Grid, theta_v_bg);
    P(Grid, exner_bg); P(Grid, exner_bg_grad); P(Grid, layer_thickness); P(Grid, trsk_indices);
    P(Grid, trsk_modified_curl_indices); P(Grid, from_index); P(Grid, to_index); P(Grid, adjacent_vector_indices_h);
    P(Grid, adjacent_signs_h); P(Grid, density_to_rhombi_indices); P(Grid, latitude_scalar); P(Grid, longitude_scalar);
    P(Grid, inner_product_weights); P(Grid, direction); P(Grid, density_to_rhombi_weights); P(Grid, trsk_weights);
    P(Grid, is_land); P(Grid, toa); P(Grid, radius); P(Grid, oro_id); P(Grid, stretching_parameter);
    P(Grid, sfc_albedo); P(Grid, sfc_rho_c); P(Grid, t_conduc_soil); P(Grid, roughness_length);
    P(Grid, latlon_interpol_indices); P(Grid, latlon_interpol_weights); P(Grid, z_soil_interface);
    P(Grid, z_soil_center); P(Grid, mean_velocity_area); P(Grid, t_const_soil); P(Grid, z_t_const);
    P(Grid, eff_hor_res);
    P(Dualgrid, area); P(Dualgrid, z_vector); P(Dualgrid, normal_distance); P(Dualgrid, from_index);
    P(Dualgrid, to_index); P(Dualgrid, vorticity_indices_triangles); P(Dualgrid, vorticity_signs_triangles);
    P(Dualgrid, f_vec);
    P(State, rho); P(State, rhotheta_v); P(State, theta_v_pert); P(State, exner_pert); P(State, wind);
    P(State, temperature_soil);
    P(Diagnostics, flux_density); P(Diagnostics, flux_density_divv); P(Diagnostics, rel_vort_on_triangles);
    P(Diagnostics, rel_vort); P(Diagnostics, pot_vort); P(Diagnostics, temperature);
    P(Diagnostics, scalar_field_placeholder); P(Diagnostics, v_squared);
    P(Forcings, pgrad_acc_old); P(Forcings, pressure_gradient_acc_neg_nl); P(Forcings, pressure_gradient_acc_neg_l);
    P(Forcings, pressure_grad_condensates_v); P(Forcings, v_squared_grad); P(Forcings, pot_vort_tend);
    P(Forcings, radiation_tendency);
    P(Irreversible_quantities, temperature_diffusion_heating); P(Irreversible_quantities, friction_acc);
    P(Irreversible_quantities, heating_diss); P(Irreversible_quantities, pressure_gradient_decel_factor);
    P(Irreversible_quantities, condensates_sediment_heat); P(Irreversible_quantities, mass_diff_tendency);
    P(Irreversible_quantities, phase_trans_rates); P(Irreversible_quantities, phase_trans_heating_rate);
    P(Config, totally_first_step_bool); P(Config, impl_thermo_weight);
    return 0;
}

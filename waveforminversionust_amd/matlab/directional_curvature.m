function [curvature, linear_seconds] = directional_curvature(state, direction, rx_index, receiver_mask) %#ok<INUSL>
    % Drop-in for Runtime/matlab/directional_curvature.m on the MI355X engine: the Gauss-Newton
    % curvature with the frozen source scales, from the forward fields the engine retained in
    % state.solver's handle (wust_oracle with nargout 3).  rx_index must be the receivers of that call.
    timer = tic;
    curvature = wust_hip_mex('curvature', state.solver.engine, double(direction), logical(receiver_mask), ...
                             state.solver.tol, state.solver.max_iter);
    linear_seconds = toc(timer);
end
